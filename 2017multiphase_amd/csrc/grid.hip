// Grid-side kernels: level set + redistancing + cell flags (a6-a8), P2G
// (a9), extrapolation and wall conditions (a10), two-phase divergence (a11),
// pressure gradient (a13), heat grid stages (a16), marching cubes (a18) and
// the solid distance/velocity fields used by the collision advection (a15).
#include "mpf_internal.h"
#include "interp.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "mc_tables.inc"

namespace mpf {

namespace {

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// a / d for the sharp kernel of the tile scatters: the hardware reciprocal
// and one Newton-Raphson correction of the quotient (Markstein's step,
// residual by fma), within an ulp of the IEEE quotient for d >= 1e-4.
__device__ __forceinline__ float div_nr(float a, float d) {
  const float y = __builtin_amdgcn_rcpf(d);
  const float q = a * y;
  const float r = fmaf(-d, q, a);
  return fmaf(r, y, q);
}

#define CELL_LOOP_HEAD                                           \
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; \
  if (idx >= g.G) return;                                        \
  int i, j, k;                                                   \
  ijk_of(g, idx, i, j, k);

// ---------------- level set (a6) ----------------
// genlevelset + mindis_cell, spray_k.cu:2753-2806.  The minimum is taken over
// squared distances and square-rooted once: sqrt is correctly rounded and
// monotone, so sqrt(min d^2) == min sqrt(d^2) bit for bit.  The 5x5x5 window
// is visited in rings of increasing lower bound on the distance from the cell
// centre to any point of a ring cell (sum over axes of max(0,|d|-1/2)^2, in
// h^2); a phase stops once its minimum is below the next ring's bound (with a
// 0.1% margin for rounding), and the 8 corner cells (bound 6.75 h^2 > 2.5^2)
// can never beat the initial 2.5h.  The result equals the full-window minimum.
// offsets (di + 4) | (dj + 4) << 8 | (dk + 4) << 16, one dword each so the
// uniform-indexed reads are scalar loads (byte tables would be vector loads
// waited on before every neighbour's address)
__constant__ int c_ring_off[120];
__constant__ int c_ring_start[10];  // ring r = offsets [start[r], start[r+1])
__constant__ float c_ring_lb[9];              // lower bound of ring r, in h^2 units

// Phases present in each 4x4x4 brick of cells (OR of the occupancy masks),
// then, per brick, in the 3x3x3 bricks around it: a superset of the phases
// in the 5x5x5 window of any cell of the brick.
__global__ void k_brick_or(Grid g, const uint8_t* __restrict__ cmask, uint8_t* __restrict__ bm, int bx, int by,
                           int bz) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)bx * by * bz) return;
  const int bk = (int)(b % bz), bj = (int)((b / bz) % by), bi = (int)(b / ((int64_t)by * bz));
  uint8_t m = 0;
  for (int di = 0; di < 4; ++di)
    for (int dj = 0; dj < 4; ++dj)
      for (int dk = 0; dk < 4; ++dk) {
        const int ci = 4 * bi + di, cj = 4 * bj + dj, ck = 4 * bk + dk;
        if (valid(g, ci, cj, ck)) m |= cmask[cid(g, ci, cj, ck)];
      }
  bm[b] = m;
}
__global__ void k_brick_dilate(const uint8_t* __restrict__ bm, uint8_t* __restrict__ dm, int bx, int by, int bz) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)bx * by * bz) return;
  const int bk = (int)(b % bz), bj = (int)((b / bz) % by), bi = (int)(b / ((int64_t)by * bz));
  uint8_t m = 0;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int i = bi + di, j = bj + dj, k = bk + dk;
        if (i >= 0 && i < bx && j >= 0 && j < by && k >= 0 && k < bz) m |= bm[((int64_t)i * by + j) * bz + k];
      }
  dm[b] = m;
}

// A phase absent from the cell's brick neighbourhood is absent from its
// window: its search is skipped, with the same result (the 2.5h cap).
__global__ void k_genlevelset(Grid g, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                              const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                              const uint8_t* __restrict__ cmask, const uint8_t* __restrict__ dm,
                              float* __restrict__ lsfluid, float* __restrict__ lsair, int8_t* __restrict__ mark) {
  const int64_t idx = xcd_block() * blockDim.x + threadIdx.x;  // XCD-contiguous cells
  if (idx >= g.G) return;
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const float h = g.h;
  mark[idx] = VACUUM;
  float r = 0.5f * h;
  f3 gp = mk3(((float)i + 0.5f) * h + g.gx0, ((float)j + 0.5f) * h + g.gy0, ((float)k + 0.5f) * h + g.gz0);
  float ma2 = INFINITY, mf2 = INFINITY;
  const float h2 = h * h;
  const int by = (g.ny + 3) / 4, bz = (g.nz + 3) / 4;
  const uint8_t near = dm[((int64_t)(i >> 2) * by + (j >> 2)) * bz + (k >> 2)];
  const bool has_f = (near & M_FLUIDSOLID) != 0, has_a = (near & (M_AIR | M_AIRSOLO)) != 0;
  for (int ring = 0; ring < 9; ++ring) {
    const float lb = 0.999f * c_ring_lb[ring] * h2;
    const bool need_f = has_f && mf2 > lb, need_a = has_a && ma2 > lb;
    if (!need_f && !need_a) break;
    const uint8_t want = (need_f ? M_FLUIDSOLID : 0) | (need_a ? (M_AIR | M_AIRSOLO) : 0);
    // the ring's cells four at a time: their masks and particle ranges are
    // read together (cells outside the grid or past the ring read cell 0
    // and are dropped), then each cell's particles are visited
    const int q1 = c_ring_start[ring + 1];
    for (int q0 = c_ring_start[ring]; q0 < q1; q0 += 4) {
      uint32_t S4[4], E4[4];
      bool take[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int q = (q0 + w < q1) ? q0 + w : q0;
        const int ow = c_ring_off[q];
        const int ci = i + (ow & 0xff) - 4, cj = j + ((ow >> 8) & 0xff) - 4, ck = k + ((ow >> 16) & 0xff) - 4;
        const bool v = q0 + w < q1 && valid(g, ci, cj, ck);
        const int64_t c = v ? cid(g, ci, cj, ck) : 0;
        const uint8_t m = cmask[c];
        S4[w] = cstart[c];
        E4[w] = cend[c];
        take[w] = v && (m & want);
      }
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        if (!take[w]) continue;
        const uint32_t s = S4[w], e = E4[w];
        // eight particles' loads issued together (the minimum does not
        // depend on the order); lanes past the range re-read the first one
        for (uint32_t p = s; p < e; p += 8) {
          f3 x[8];
          int8_t fl[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const bool in = p + u < e;
            x[u] = pos[in ? p + u : s];
            fl[u] = in ? flag[p + u] : (int8_t)-1;
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            f3 d = sub3(x[u], gp);
            float d2 = dot3(d, d);
            const int8_t f = fl[u];
            if (f == AIR || f == AIRSOLO) ma2 = fminf(ma2, d2);
            else if (f == FLUID || f == SOLID) mf2 = fminf(mf2, d2);
          }
        }
      }
    }
  }
  float mair = 2.5f * h, mfl = 2.5f * h;
  float da = sqrtf(ma2), df = sqrtf(mf2);
  mair = (da < mair) ? da : mair;
  mfl = (df < mfl) ? df : mfl;
  lsfluid[idx] = mfl - r;
  lsair[idx] = mair - r;
}

// markLS_bigpositive spray_k.cu:3683-3697 (mode 0) / preparels 3708-3722 (mode 1)
__global__ void k_ls_prepare(int64_t G, const float* __restrict__ ls, float* __restrict__ out,
                             int8_t* __restrict__ mark, float h, int mode) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= G) return;
  float v = ls[idx];
  if (mode == 0) {
    v = v / h;
    if (v > 1.99f) { v = 5.0f; mark[idx] = AIR; }
    else mark[idx] = FLUID;
  } else {
    v = -v / h;
    if (v > 0) { v = 5.0f; mark[idx] = AIR; }
    else mark[idx] = FLUID;
  }
  out[idx] = v;
}
// setLSback_bigpositive spray_k.cu:3699-3706 (mode 0) / setLSback 3724-3731 (mode 1)
__global__ void k_ls_back(int64_t G, float* __restrict__ ls, float h, int mode) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= G) return;
  ls[idx] = (mode == 0) ? ls[idx] * h : -ls[idx] * h;
}

// One triple of solvedistance (spray_k.cu:746-754: d = min(a,b,c) + 1, or
// (a + b + c + sqrt(3 - (a-b)^2 - (a-c)^2 - (b-c)^2)) / 3 when that d exceeds
// max(a,b,c); x = min(x, d)) with the division of its three-axis branch left
// out: those candidates enter x only through a minimum and x/3 rounded
// to nearest is non-decreasing in x, so the smallest numerator s divided once
// at the end (solvedistance_end) gives the same float as dividing each.
__device__ __forceinline__ void solvedistance_acc(float a, float b, float c, float& x, float& s) {
  const float d = fminf(a, fminf(b, c)) + 1.0f;
  if (d > fmaxf(a, fmaxf(b, c)))
    s = fminf(s, a + b + c + sqrtf(3.0f - (a - b) * (a - b) - (a - c) * (a - c) - (b - c) * (b - c)));
  else if (d < x)
    x = d;
}
__device__ __forceinline__ void solvedistance_end(float s, float& x) {
  const float d = s / 3.0f;  // infinity when no triple took the branch
  if (d < x) x = d;
}
// sweepphibytype (spray_k.cu:771-789) / sweepphi (755-770, typeflag < 0), two
// buffers, on 1 x SW_J x SW_K cell tiles of one wave's cells (lane = row *
// SW_K + column), with change flags.  A Jacobi pass changes a cell only when
// a value of its 7-point neighbourhood changed in the pass before.  cur[tile]
// records which cells of the tile got a smaller value: SF_ANY for any, and
// the j-low, j-high, k-low and k-high bits for a change on that face of the
// tile.  A tile whose own flags, those of its x neighbours and the facing bits
// of its y and z neighbours of the previous pass (prev) are clear would
// recompute its previous values, which the output buffer (the previous pass's
// input) already holds, so it is skipped and gets a clear flag.  The first
// pass of a series (prev == nullptr) computes every tile.  A wave owns
// SW_L consecutive tiles: its first SW_L lanes read their flags together and
// the wave then visits the active ones.
constexpr uint8_t SF_ANY = 1, SF_JLO = 2, SF_JHI = 4, SF_KLO = 8, SF_KHI = 16;
constexpr int SW_L = 16;
constexpr int SW_U1 = 4;
static_assert(SW_J * SW_K == 64 && SW_L <= 64, "one wave's cells per eikonal tile");
// whether tile t (of a ntx x ntj x ntk tile grid, x slowest) must be computed:
// always without previous flags, else when it or an x neighbour changed or a
// y or z neighbour changed on the facing side (the seven flags are read
// together: indices clamped, then masked)
__device__ __forceinline__ bool tile_active(const uint8_t* __restrict__ prev, uint32_t t, bool mine, uint32_t ntj,
                                            uint32_t ntk, uint32_t ntx) {
  if (!mine) return false;
  if (!prev) return true;
  const uint32_t sxt = ntj * ntk;
  const uint32_t tk = t % ntk, tj = (t / ntk) % ntj, i = t / sxt;
  const uint8_t f0 = prev[t];
  const uint8_t fx0 = prev[i > 0 ? t - sxt : t], fx1 = prev[i + 1 < ntx ? t + sxt : t];
  const uint8_t fy0 = prev[tj > 0 ? t - ntk : t], fy1 = prev[tj + 1 < ntj ? t + ntk : t];
  const uint8_t fz0 = prev[tk > 0 ? t - 1 : t], fz1 = prev[tk + 1 < ntk ? t + 1 : t];
  const int a = (f0 & SF_ANY) | ((i > 0) ? (fx0 & SF_ANY) : 0) | ((i + 1 < ntx) ? (fx1 & SF_ANY) : 0) |
                ((tj > 0) ? (fy0 & SF_JHI) : 0) | ((tj + 1 < ntj) ? (fy1 & SF_JLO) : 0) |
                ((tk > 0) ? (fz0 & SF_KHI) : 0) | ((tk + 1 < ntk) ? (fz1 & SF_KLO) : 0);
  return a != 0;
}
// the flags of a tile from its lanes' change bits (lane = row * SW_K + column)
__device__ __forceinline__ uint8_t tile_flags(bool changed) {
  const uint64_t bal = __ballot(changed);
  constexpr uint64_t JLO = (1ull << SW_K) - 1, JHI = JLO << (64 - SW_K);
  constexpr uint64_t KLO = 0x0001000100010001ull, KHI = KLO << (SW_K - 1);
  static_assert(SW_K == 16 && SW_J == 4, "face masks of a 4 x 16 tile");
  return (uint8_t)((bal ? SF_ANY : 0) | ((bal & JLO) ? SF_JLO : 0) | ((bal & JHI) ? SF_JHI : 0) |
                   ((bal & KLO) ? SF_KLO : 0) | ((bal & KHI) ? SF_KHI : 0));
}
__global__ void __launch_bounds__(256)
    k_sweep(Grid g, const float* __restrict__ in, float* __restrict__ out, const int8_t* __restrict__ mark,
            int typeflag, const uint8_t* __restrict__ prev, uint8_t* __restrict__ cur, uint32_t ntj, uint32_t ntk,
            uint32_t ntiles) {
  const uint32_t t0 = (blockIdx.x * 4u + threadIdx.x / 64u) * SW_L;
  if (t0 >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const uint32_t sxt = ntj * ntk;
  const uint32_t tl = t0 + lane;
  const bool mine = lane < SW_L && tl < ntiles;
  const bool act = tile_active(prev, tl, mine, ntj, ntk, (uint32_t)g.nx);
  uint64_t todo = __ballot(act);
  uint8_t myflag = 0;
  const int jl = lane / SW_K, kl = lane % SW_K;
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  while (todo) {
    const int q = __builtin_ctzll(todo);
    todo &= todo - 1;
    const uint32_t t = t0 + q;
    const int tk = (int)(t % ntk), tj = (int)((t / ntk) % ntj), i = (int)(t / sxt);
    const int j = tj * SW_J + jl, k = tk * SW_K + kl;
    bool changed = false;
    if (j < g.ny && k < g.nz) {
      const int64_t idx = ((int64_t)i * g.ny + j) * g.nz + k;
      // the cell, its mark and its six neighbours (x-, x+, y-, y+, z-, z+)
      // are loaded together (a neighbour outside the grid reads the cell
      // itself and is replaced by infinity)
      const bool vx0 = i > 0, vx1 = i < g.nx - 1, vy0 = j > 0, vy1 = j < g.ny - 1, vz0 = k > 0, vz1 = k < g.nz - 1;
      const float c0 = in[idx];
      float res = c0;
      const int8_t m = (typeflag >= 0) ? mark[idx] : (int8_t)0;
      const float lx0 = in[vx0 ? idx - sx : idx], lx1 = in[vx1 ? idx + sx : idx];
      const float ly0 = in[vy0 ? idx - sy : idx], ly1 = in[vy1 ? idx + sy : idx];
      const float lz0 = in[vz0 ? idx - 1 : idx], lz1 = in[vz1 ? idx + 1 : idx];
      if (!(typeflag >= 0 && m == typeflag)) {
        const float nx0 = vx0 ? lx0 : INFINITY, nx1 = vx1 ? lx1 : INFINITY;
        const float ny0 = vy0 ? ly0 : INFINITY, ny1 = vy1 ? ly1 : INFINITY;
        const float nz0 = vz0 ? lz0 : INFINITY, nz1 = vz1 ? lz1 : INFINITY;
        // an update needs a neighbour below the current value (every
        // candidate is at least min(a, b, c)), so cells without one keep
        // their value without evaluating the 8 triples
        const float mn = fminf(fminf(fminf(nx0, nx1), fminf(ny0, ny1)), fminf(nz0, nz1));
        if (mn < res) {
          float s3 = INFINITY;
          for (int di = -1; di <= 1; di += 2)
            for (int dj = -1; dj <= 1; dj += 2)
              for (int dk = -1; dk <= 1; dk += 2)
                if ((di < 0 ? vx0 : vx1) && (dj < 0 ? vy0 : vy1) && (dk < 0 ? vz0 : vz1))
                  solvedistance_acc(di < 0 ? nx0 : nx1, dj < 0 ? ny0 : ny1, dk < 0 ? nz0 : nz1, res, s3);
          solvedistance_end(s3, res);
        }
      }
      // res only ever takes strictly smaller values: unchanged means bitwise equal
      changed = res < c0;
      out[idx] = res;
    }
    const uint8_t f = tile_flags(changed);
    if (lane == q) myflag = f;
  }
  if (mine) cur[tl] = myflag;
}

// mergeLSAndMarkGrid spray_k.cu:3733-3777
__global__ void k_merge(Grid g, const float* __restrict__ lsfluid, const float* __restrict__ lsair,
                        float* __restrict__ lsmerge, int8_t* __restrict__ mark) {
  CELL_LOOP_HEAD
  const float h = g.h;
  float lf = lsfluid[idx], la = lsair[idx], lm;
  int8_t m;
  if (la > 4.99f * h) { lm = lf; m = (lf > 0) ? VACUUM : FLUID; }
  else if (lf > 4.99f * h) { lm = la; m = (la > 0) ? VACUUM : AIR; }
  else if (la > 0.8f * h && lf > 0.8f * h) { m = VACUUM; lm = fminf(lf, la); }
  else { lm = (lf - la) * 0.5f; m = (lm > 0) ? AIR : FLUID; }
  if (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1 || k == 0 || k == g.nz - 1) {
    m = BOUNDARY;
    lm = -0.5f * h;
  }
  lsmerge[idx] = lm;
  mark[idx] = m;
}

// computePhigra spray_k.cu:3994-4019
__global__ void k_phigrad(Grid g, const float* __restrict__ ls, float* __restrict__ gx, float* __restrict__ gy,
                          float* __restrict__ gz) {
  CELL_LOOP_HEAD
  float c0 = ls[idx];
  float x1 = valid(g, i + 1, j, k) ? ls[cid(g, i + 1, j, k)] : c0;
  float x0 = valid(g, i - 1, j, k) ? ls[cid(g, i - 1, j, k)] : c0;
  float y1 = valid(g, i, j + 1, k) ? ls[cid(g, i, j + 1, k)] : c0;
  float y0 = valid(g, i, j - 1, k) ? ls[cid(g, i, j - 1, k)] : c0;
  float z1 = valid(g, i, j, k + 1) ? ls[cid(g, i, j, k + 1)] : c0;
  float z0 = valid(g, i, j, k - 1) ? ls[cid(g, i, j, k - 1)] : c0;
  gx[idx] = ((x1 - x0) * 0.5f) / g.h;
  gy[idx] = ((y1 - y0) * 0.5f) / g.h;
  gz[idx] = ((z1 - z0) * 0.5f) / g.h;
}

// markfluid_dense + markBoundaryCell spray_k.cu:506-545 (all cells)
__global__ void k_mark_counts(Grid g, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                              const int8_t* __restrict__ flag, int8_t* __restrict__ mark) {
  CELL_LOOP_HEAD
  int cfs = 0, ca = 0;
  uint32_t s = cstart[idx];
  if (s != MPF_CELL_UNDEF)
    for (uint32_t p = s; p < cend[idx]; ++p) {
      int8_t f = flag[p];
      if (f == FLUID || f == SOLID) cfs++;
      else if (f == AIR) ca++;
    }
  int8_t m = (cfs == 0 && ca == 0) ? VACUUM : (cfs > ca ? FLUID : AIR);
  if (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1 || k == 0 || k == g.nz - 1) m = BOUNDARY;
  mark[idx] = m;
}

// Solid constraint: non-boundary cells whose particles are at least half SOLID.
__global__ void k_mark_solid(Grid g, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                             const int8_t* __restrict__ flag, const uint8_t* __restrict__ cmask,
                             int8_t* __restrict__ mark) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  if (!(cmask[idx] & M_SOLID) || mark[idx] == BOUNDARY) return;
  int cs = 0, co = 0;
  for (uint32_t p = cstart[idx]; p < cend[idx]; ++p) {
    if (flag[p] == SOLID) cs++;
    else co++;
  }
  if (cs > 0 && cs >= co) mark[idx] = SOLID;
}

// ---------------- P2G (a9) ----------------
// mapvelp2g_k_fluidSolid / mapvelp2g_k_air with sumcell_* and sharp_kernel,
// spray_k.cu:255-258, 2809-2937.  One thread per face of axis A; windows
// normal [-1,0] x transverse [-1,1] (liquid+solid) and normal [-2,1] x
// transverse [-2,2] (air); cells without particles of the phase are skipped
// through the occupancy mask.
template <int A, bool AIRPH>
__global__ void k_p2g(Grid g, const f3* __restrict__ pos, const f3* __restrict__ vel, const float* __restrict__ mass,
                      const int8_t* __restrict__ flag, const uint32_t* __restrict__ cstart,
                      const uint32_t* __restrict__ cend, const uint8_t* __restrict__ cmask,
                      const uint8_t* __restrict__ dm, float* __restrict__ out, float* __restrict__ out2,
                      int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  constexpr uint8_t need = AIRPH ? M_AIR : M_FLUIDSOLID;
  {
    // the window (offsets -2..2 from cell (i,j,k)) lies in the 3x3x3 bricks
    // around that cell's brick: no particle of the phase there, no weight
    const int by = (g.ny + 3) / 4, bz = (g.nz + 3) / 4;
    const int64_t b = ((int64_t)(min(i, g.nx - 1) >> 2) * by + (min(j, g.ny - 1) >> 2)) * bz + (min(k, g.nz - 1) >> 2);
    if (!(dm[b] & need)) {
      out[f] = 0.0f;
      if (out2) out2[f] = 0.0f;
      return;
    }
  }
  const float scale = 1.0f / g.h;
  f3 gp = mk3((float)i + (A == 0 ? 0.0f : 0.5f), (float)j + (A == 1 ? 0.0f : 0.5f), (float)k + (A == 2 ? 0.0f : 0.5f));
  constexpr int lo = AIRPH ? -2 : -1, hin = AIRPH ? 1 : 0, lot = AIRPH ? -2 : -1, hit = AIRPH ? 2 : 1;
  float weight = 0, usum = 0;
  const float RE = 1.4f;
  const float RE2 = RE * RE;
  for (int di = (A == 0 ? lo : lot); di <= (A == 0 ? hin : hit); ++di)
    for (int dj = (A == 1 ? lo : lot); dj <= (A == 1 ? hin : hit); ++dj)
      for (int dk = (A == 2 ? lo : lot); dk <= (A == 2 ? hin : hit); ++dk) {
        if (!valid(g, i + di, j + dj, k + dk)) continue;
        int64_t c = cid(g, i + di, j + dj, k + dk);
        if (!(cmask[c] & need)) continue;
        uint32_t s = cstart[c], e = cend[c];
        for (uint32_t p = s; p < e; ++p) {
          int8_t fl = flag[p];
          bool take = AIRPH ? (fl == AIR) : (fl == FLUID || fl == SOLID);
          if (!take) continue;
          f3 x = pos[p];
          f3 dv = mk3((x.x - g.gx0) * scale - gp.x, (x.y - g.gy0) * scale - gp.y, (x.z - g.gz0) * scale - gp.z);
          float r2 = dot3(dv, dv);
          float w = mass[p] * fmaxf(RE2 / fmaxf(r2, 0.0001f) - 1.0f, 0.0f);
          weight += w;
          const float* vp = &vel[p].x;
          usum += w * vp[A];
        }
      }
  const float u = (weight > 0) ? (usum / weight) : 0.0f;
  out[f] = u;
  if (out2) out2[f] = u;
}

// Liquid+solid P2G, tile scatter (the hot path).  A block owns a 4x8x8 cell
// tile with one thread per cell and the cell's particles in registers.  Its
// cells contribute to the faces of a per-axis region (the tile's faces plus
// the ring its windows reach); the 27 window offsets are visited as phases in
// the gather's lexicographic order, and in phase o every thread adds its
// particles, in index order, to face c - o of each axis whose window holds o
// -- one-to-one within a phase, so LDS accumulates without atomics.  Each
// tile stores its region's (sum w, sum w v); a second kernel adds, per face,
// the regions of the (at most 8) tiles that cover it in tile order and
// divides.  Every particle-face pair is evaluated once and every thread holds
// a tile cell (no halo threads); sums are deterministic, associated per tile
// (a face inside one tile's region sums exactly like the gather).
constexpr int ST_X = 4, ST_Y = 8, ST_Z = 8;
constexpr int ST_CELLS = ST_X * ST_Y * ST_Z;  // 256 threads
__host__ __device__ constexpr int st_t(int d) { return d == 0 ? ST_X : (d == 1 ? ST_Y : ST_Z); }
// region of axis a: faces [t0_d - (d != a), t0_d + T_d] along d
__host__ __device__ constexpr int st_r(int a, int d) { return st_t(d) + (d == a ? 1 : 2); }
__host__ __device__ constexpr int st_rsz(int a) { return st_r(a, 0) * st_r(a, 1) * st_r(a, 2); }
__host__ __device__ constexpr int st_roff(int a) { return a == 0 ? 0 : (a == 1 ? st_rsz(0) : st_rsz(0) + st_rsz(1)); }
constexpr int ST_RTOT = st_rsz(0) + st_rsz(1) + st_rsz(2);  // 1580 faces
constexpr int SCH = 8;   // particles of a cell in registers in the first pass
constexpr int SCH2 = 4;  // particles per cell in each overflow pass

// A cell's next ns particles [base, base + cnt) in registers (an empty slot
// loads the chunk's last particle, or particle 0 when the chunk is empty, so
// all loads are in flight together; its weight is 0).
template <int NS>
__device__ __forceinline__ void p2g_load(const f3* __restrict__ pos, const f3* __restrict__ vel,
                                         const float* __restrict__ mass, const int8_t* __restrict__ flag,
                                         const Grid& g, float scale, uint32_t base, uint32_t cnt, float (&X)[NS],
                                         float (&Y)[NS], float (&Z)[NS], float (&VX)[NS], float (&VY)[NS],
                                         float (&VZ)[NS], float (&M)[NS]) {
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const uint32_t p = (cnt > 0) ? base + min((uint32_t)q, cnt - 1) : 0u;
    const int8_t fl = flag[p];
    const f3 x = pos[p], v = vel[p];
    const float m = mass[p];
    X[q] = (x.x - g.gx0) * scale;
    Y[q] = (x.y - g.gy0) * scale;
    Z[q] = (x.z - g.gz0) * scale;
    VX[q] = v.x; VY[q] = v.y; VZ[q] = v.z;
    M[q] = ((uint32_t)q < cnt && (fl == FLUID || fl == SOLID)) ? m : 0.0f;
  }
}

// The 27 phases of one pass: in phase o every active thread adds its NS
// register particles (of cell (ci, cj, ck)) to face c - o of each axis whose
// window holds o.  Within a pass the active cells are distinct, so a face
// takes at most one thread per phase (no atomics).  Inactive threads only
// meet the barriers.
template <int NS>
__device__ __forceinline__ void p2g_phases(int x0, int y0, int z0, int ci, int cj, int ck, bool act,
                                           const float (&X)[NS], const float (&Y)[NS], const float (&Z)[NS],
                                           const float (&VX)[NS], const float (&VY)[NS], const float (&VZ)[NS],
                                           const float (&M)[NS], float* sW, float* sU) {
  const float RE = 1.4f, RE2 = RE * RE;
  for (int ph = 0; ph < 27; ++ph) {
    const int o[3] = {ph / 9 - 1, (ph / 3) % 3 - 1, ph % 3 - 1};
    if (act) {
      // The faces of the three axes in this phase share the cell c - o:
      // axis a's face sits at that cell's corner along a and at its centre
      // along the others, so each particle's six coordinate differences
      // (corner and centre, per direction) and their squares serve all
      // three distances, summed in the gather's order.
      const int f[3] = {ci - o[0], cj - o[1], ck - o[2]};
      int li[3];
      float W[3], U[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int r0 = x0 - (a != 0), r1 = y0 - (a != 1), r2 = z0 - (a != 2);
        li[a] = st_roff(a) + ((f[0] - r0) * st_r(a, 1) + (f[1] - r1)) * st_r(a, 2) + (f[2] - r2);
        W[a] = (o[a] > 0) ? 0.0f : sW[li[a]];  // normal window is {-1, 0}
        U[a] = (o[a] > 0) ? 0.0f : sU[li[a]];
      }
      const float fx = (float)f[0], fy = (float)f[1], fz = (float)f[2];
      const float hx = fx + 0.5f, hy = fy + 0.5f, hz = fz + 0.5f;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const float ax = X[q] - fx, bx = X[q] - hx, ay = Y[q] - fy, by = Y[q] - hy, az = Z[q] - fz, bz = Z[q] - hz;
        const float ax2 = ax * ax, bx2 = bx * bx, ay2 = ay * ay, by2 = by * by, az2 = az * az, bz2 = bz * bz;
        const float d2[3] = {ax2 + by2 + bz2, bx2 + ay2 + bz2, bx2 + by2 + az2};
        const float vq[3] = {VX[q], VY[q], VZ[q]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (o[a] > 0) continue;
          const float w = M[q] * fmaxf(div_nr(RE2, fmaxf(d2[a], 0.0001f)) - 1.0f, 0.0f);
          W[a] += w;
          U[a] += w * vq[a];
        }
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (o[a] > 0) continue;
        sW[li[a]] = W[a];
        sU[li[a]] = U[a];
      }
    }
    __syncthreads();
  }
}

// Overflow passes of a tile: pass p takes particles [thr, thr + SCH2) of
// every cell holding more than thr = SCH + (p-1) SCH2, the cells compacted
// onto the first threads (a wave without such a cell skips the arithmetic).
// Returns after the pass that finds no cell left; block-uniform.
template <class PassFn>
__device__ __forceinline__ void overflow_passes(uint32_t n, uint32_t* sList, int* sWc, PassFn pass) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (uint32_t thr = SCH;; thr += SCH2) {
    const bool pred = n > thr;
    const unsigned long long bal = __ballot(pred);
    if (ln == 0) sWc[wv] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w2 = 0; w2 < ST_CELLS / 64; ++w2) {
      off += (w2 < wv) ? sWc[w2] : 0;
      tot += sWc[w2];
    }
    if (pred) sList[off + __popcll(bal & ((1ull << ln) - 1ull))] = threadIdx.x;
    __syncthreads();
    if (tot == 0) return;
    const bool act = (int)threadIdx.x < tot;
    pass(act, act ? sList[threadIdx.x] : 0u, thr);
  }
}

__global__ void __launch_bounds__(ST_CELLS)
    k_p2g_tile(Grid g, const f3* __restrict__ pos, const f3* __restrict__ vel, const float* __restrict__ mass,
               const int8_t* __restrict__ flag, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
               const uint8_t* __restrict__ cmask, float* __restrict__ parts, int ntx, int nty, int ntz,
               const float* __restrict__ band, float big, const uint8_t* __restrict__ dmask) {
  __shared__ float sW[ST_RTOT];
  __shared__ float sU[ST_RTOT];
  __shared__ uint32_t sS[ST_CELLS], sN[ST_CELLS], sList[ST_CELLS];
  __shared__ int sWc[ST_CELLS / 64];
  const int tl = (int)xcd_block();  // XCD-contiguous tiles
  int t = tl;
  const int tz = t % ntz;
  t /= ntz;
  const int ty = t % nty;
  const int tx = t / nty;
  const int x0 = tx * ST_X, y0 = ty * ST_Y, z0 = tz * ST_Z;
  for (int q = threadIdx.x; q < ST_RTOT; q += ST_CELLS) {
    sW[q] = 0.0f;
    sU[q] = 0.0f;
  }
  const int b = threadIdx.x;
  const int ci = x0 + b / (ST_Y * ST_Z), cj = y0 + (b / ST_Z) % ST_Y, ck = z0 + b % ST_Z;
  uint32_t s = 0, e = 0;
  if (valid(g, ci, cj, ck)) {
    const int64_t c = cid(g, ci, cj, ck);
    if (cmask[c] & M_FLUIDSOLID) {
      s = cstart[c];
      e = cend[c];
    }
  }
  // band mode: a tile whose cells and ring hold no cell with band < big maps nothing
  bool near = true;
  if (band) {
    // band cells lie within 3 cells of a solid cell, and the region's cells
    // within one cell of the tile, so a tile whose bricks (1 x 2 x 2 of the
    // 4x4x4 bricks, each with its 3x3x3 neighbourhood in dmask) hold no solid
    // cell has none in its region: its band scan is skipped
    static_assert(ST_X == 4 && ST_Y == 8 && ST_Z == 8, "a tile spans 1 x 2 x 2 bricks");
    const int nby = (g.ny + 3) / 4, nbz = (g.nz + 3) / 4;
    uint8_t bm = 0;
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int dk = 0; dk < 2; ++dk) {
        const int bj = (y0 >> 2) + dj, bk = (z0 >> 2) + dk;
        if (bj < nby && bk < nbz) bm |= dmask[((int64_t)(x0 >> 2) * nby + bj) * nbz + bk];
      }
    near = (bm & M_SOLID) != 0;
  }
  if (band && near) {
    bool mine = false;
    for (int q = threadIdx.x; q < (ST_X + 2) * (ST_Y + 2) * (ST_Z + 2); q += ST_CELLS) {
      const int bi = x0 - 1 + q / ((ST_Y + 2) * (ST_Z + 2)), bj = y0 - 1 + (q / (ST_Z + 2)) % (ST_Y + 2),
                bk = z0 - 1 + q % (ST_Z + 2);
      if (valid(g, bi, bj, bk) && band[cid(g, bi, bj, bk)] < big) mine = true;
    }
    near = __syncthreads_or(mine);
  }
  const int any = __syncthreads_or(e > s) && near;
  if (any) {
    const float scale = 1.0f / g.h;
    const uint32_t n = e - s;
    sS[b] = s;
    sN[b] = n;
    {
      float X[SCH], Y[SCH], Z[SCH], VX[SCH], VY[SCH], VZ[SCH], M[SCH];
      p2g_load<SCH>(pos, vel, mass, flag, g, scale, s, min(n, (uint32_t)SCH), X, Y, Z, VX, VY, VZ, M);
      p2g_phases<SCH>(x0, y0, z0, ci, cj, ck, n > 0, X, Y, Z, VX, VY, VZ, M, sW, sU);
    }
    // cells with more than SCH particles: the rest in compacted passes
    overflow_passes(n, sList, sWc, [&](bool act, uint32_t cb, uint32_t thr) {
      const int oi = x0 + (int)cb / (ST_Y * ST_Z), oj = y0 + ((int)cb / ST_Z) % ST_Y, ok = z0 + (int)cb % ST_Z;
      const uint32_t os = act ? sS[cb] + thr : 0u;
      const uint32_t oc = act ? min(sN[cb] - thr, (uint32_t)SCH2) : 0u;
      float X[SCH2], Y[SCH2], Z[SCH2], VX[SCH2], VY[SCH2], VZ[SCH2], M[SCH2];
      if (act) p2g_load<SCH2>(pos, vel, mass, flag, g, scale, os, oc, X, Y, Z, VX, VY, VZ, M);
      p2g_phases<SCH2>(x0, y0, z0, oi, oj, ok, act, X, Y, Z, VX, VY, VZ, M, sW, sU);
    });
  } else {
    __syncthreads();
  }
  float* dst = parts + (int64_t)tl * (2 * ST_RTOT);
  for (int q = threadIdx.x; q < ST_RTOT; q += ST_CELLS) {
    dst[q] = sW[q];
    dst[ST_RTOT + q] = sU[q];
  }
}

// u = (sum w v) / (sum w) per face from the tile regions covering it, added in
// tile order (also into out2 when given: the copy the FLIP update keeps).
// Band mode: faces without a cell of band < big beside them are 0.
template <int A>
__global__ void k_p2g_collect(Grid g, const float* __restrict__ parts, int ntx, int nty, int ntz,
                              float* __restrict__ out, float* __restrict__ out2, int64_t nf,
                              const float* __restrict__ band, float big) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  if (band) {
    const int n3a = (A == 0) ? g.nx : (A == 1) ? g.ny : g.nz;
    const int na = (A == 0) ? i : (A == 1) ? j : k;
    bool in_band = false;
    if (na < n3a) in_band = in_band || band[cid(g, i, j, k)] < big;
    if (na > 0) in_band = in_band || band[cid(g, i - (A == 0), j - (A == 1), k - (A == 2))] < big;
    if (!in_band) {
      out[f] = 0.0f;
      return;
    }
  }
  const int fv[3] = {i, j, k};
  const int nt[3] = {ntx, nty, ntz};
  int qlo[3], qhi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // tiles q with q*T - (d != A) <= f <= q*T + T
    const int T = st_t(d);
    int lo = (fv[d] - T + T - 1) / T;  // ceil((f - T) / T) for f >= 0
    if (fv[d] - T < 0) lo = 0;
    int hi = (fv[d] + (d != A ? 1 : 0)) / T;
    if (hi > nt[d] - 1) hi = nt[d] - 1;
    qlo[d] = lo;
    qhi[d] = hi;
  }
  float W = 0.0f, U = 0.0f;
  for (int qx = qlo[0]; qx <= qhi[0]; ++qx)
    for (int qy = qlo[1]; qy <= qhi[1]; ++qy)
      for (int qz = qlo[2]; qz <= qhi[2]; ++qz) {
        const int r0 = qx * ST_X - (A != 0), r1 = qy * ST_Y - (A != 1), r2 = qz * ST_Z - (A != 2);
        const int li = st_roff(A) + ((fv[0] - r0) * st_r(A, 1) + (fv[1] - r1)) * st_r(A, 2) + (fv[2] - r2);
        const float* src = parts + ((int64_t)(qx * nty + qy) * ntz + qz) * (2 * ST_RTOT);
        W += src[li];
        U += src[ST_RTOT + li];
      }
  const float u = (W > 0) ? (U / W) : 0.0f;
  out[f] = u;
  if (out2) out2[f] = u;
}

// ---------------- extrapolation + wall conditions (a10) ----------------
// initphi, spray_k.cu:712-723
__global__ void k_initphi(int64_t G, const int8_t* __restrict__ mark, float* __restrict__ phi, int8_t type, float big) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < G) phi[idx] = (mark[idx] == type) ? -0.5f : big;
}

// sweepu_k_bubble spray_k.cu:3779-3880 (one axis, out-of-place): the new
// value of face (i, j, k) of axis A (index f); u0 returns its old value
template <int A>
__device__ __forceinline__ float sweep_u_face(const Grid& g, const float* __restrict__ u,
                                              const float* __restrict__ ls, const int8_t* __restrict__ mark, int8_t sf,
                                              int i, int j, int k, int64_t f, float& u0) {
  const int n[3] = {i, j, k};
  const int lim = (A == 0) ? g.nx : (A == 1) ? g.ny : g.nz;
  const bool inner = n[A] > 1 && n[A] < lim - 1;
  // first round of loads: the face, the marks and level sets of its two cells
  // (cell indices clamped for the faces the rule leaves as they are)
  const int ci = min(i, g.nx - 1), cj = min(j, g.ny - 1), ck = min(k, g.nz - 1);
  const int bi = max(i - (A == 0), 0), bj = max(j - (A == 1), 0), bk = max(k - (A == 2), 0);
  float res = u[f];
  u0 = res;
  const int8_t mc = mark[cid(g, ci, cj, ck)], mb = mark[cid(g, bi, bj, bk)];
  const float lsc = ls[cid(g, ci, cj, ck)], lsb = ls[cid(g, bi, bj, bk)];
  if (inner && mc != sf && mb != sf) {
    constexpr int t1 = (A == 0) ? 1 : (A == 1) ? 0 : 1;
    constexpr int t2 = (A == 0) ? 2 : (A == 1) ? 2 : 0;
    const int lims[3] = {g.nx, g.ny, g.nz};
    // second round: the level sets beside both cells across each transverse
    // axis and the face's six neighbours, all issued together (indices
    // clamped; a value outside the grid is not used)
    float lt[2][2][2];  // [transverse t][sign][cell c / cell b]
    bool in[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int ax = t == 0 ? t1 : t2;
#pragma unroll
      for (int sg = 0; sg < 2; ++sg) {
        const int dd = sg ? 1 : -1;
        in[t][sg] = n[ax] + dd >= 0 && n[ax] + dd <= lims[ax] - 1;
        int o[3] = {0, 0, 0};
        o[ax] = in[t][sg] ? dd : 0;
        lt[t][sg][0] = ls[cid(g, i + o[0], j + o[1], k + o[2])];
        lt[t][sg][1] = ls[cid(g, i - (A == 0) + o[0], j - (A == 1) + o[1], k - (A == 2) + o[2])];
      }
    }
    float un[3][2];  // u of the face's neighbours along x, y, z (sign -1, +1)
    const int fl[3] = {(A == 0) ? g.nx : g.nx - 1, (A == 1) ? g.ny : g.ny - 1, (A == 2) ? g.nz : g.nz - 1};
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int sg = 0; sg < 2; ++sg) {
        int q[3] = {i, j, k};
        q[d] = min(max(q[d] + (sg ? 1 : -1), 0), fl[d]);
        un[d][sg] = u[fid(g, A, q[0], q[1], q[2])];
      }
    // The reference visits the 8 octants (di, dj, dk in {-1, +1}, di
    // outermost) and the last one whose three weights are all >= 0 sets the
    // value.  Each weight depends on its own axis' sign only, so that octant
    // takes, per axis, +1 if its weight passes there, else -1.
    const float dl = lsc - lsb;
    const float wa[2] = {(float)(1) * dl, (float)(-1) * dl};  // sign -1, +1
    float wt[2][2];
    bool ok[3][2];
    ok[A][0] = !(wa[0] < 0);
    ok[A][1] = !(wa[1] < 0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int ax = t == 0 ? t1 : t2;
#pragma unroll
      for (int sg = 0; sg < 2; ++sg) {
        float w = 0.0f;
        if (in[t][sg]) w = (lsc + lsb - lt[t][sg][0] - lt[t][sg][1]) * 0.5f;
        wt[t][sg] = w;
        ok[ax][sg] = in[t][sg] && !(w < 0);
      }
    }
    if ((ok[0][0] || ok[0][1]) && (ok[1][0] || ok[1][1]) && (ok[2][0] || ok[2][1])) {
      const int sgn[3] = {ok[0][1] ? 1 : 0, ok[1][1] ? 1 : 0, ok[2][1] ? 1 : 0};
      float wv[3];
      wv[A] = wa[sgn[A]];
      wv[t1] = wt[0][sgn[t1]];
      wv[t2] = wt[1][sgn[t2]];
      float wx = wv[0], wy = wv[1], wz = wv[2];
      float wsum = wx + wy + wz;
      if (wsum == 0) wx = wy = wz = 1.0f / 3;
      else { wx /= wsum; wy /= wsum; wz /= wsum; }
      res = wx * un[0][sgn[0]] + wy * un[1][sgn[1]] + wz * un[2][sgn[2]];
    }
  }
  return res;
}

// One extrapolation pass of axis A on 1 x SW_J x SW_K tiles of its faces,
// with the change flags of the eikonal passes: a face's new value reads its
// own and its six neighbours' values (level sets and marks are fixed over the
// passes), so a tile whose neighbourhood did not change in the previous pass
// (prev) is skipped, its values being in the output buffer already.  cur (may
// be null) records the changes of this pass; prev == nullptr computes every
// tile.  A change is a different bit pattern.
template <int A>
__global__ void __launch_bounds__(256)
    k_sweep_u(Grid g, const float* __restrict__ u, float* __restrict__ out, const float* __restrict__ ls,
              const int8_t* __restrict__ mark, int8_t sf, const uint8_t* __restrict__ prev, uint8_t* __restrict__ cur,
              uint32_t ntj, uint32_t ntk, uint32_t ntiles, int per_wave) {
  const uint32_t t0 = (blockIdx.x * 4u + threadIdx.x / 64u) * (uint32_t)per_wave;
  if (t0 >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int nfx = g.nx + (A == 0), nfy = g.ny + (A == 1), nfz = g.nz + (A == 2);
  const bool act = tile_active(prev, t0 + lane, lane < per_wave && t0 + lane < ntiles, ntj, ntk, (uint32_t)nfx);
  uint64_t todo = __ballot(act);
  uint8_t myflag = 0;
  const int jl = lane / SW_K, kl = lane % SW_K;
  const uint32_t sxt = ntj * ntk;
  while (todo) {
    const int q = __builtin_ctzll(todo);
    todo &= todo - 1;
    const uint32_t t = t0 + q;
    const int tk = (int)(t % ntk), tj = (int)((t / ntk) % ntj), i = (int)(t / sxt);
    const int j = tj * SW_J + jl, k = tk * SW_K + kl;
    bool changed = false;
    if (j < nfy && k < nfz) {
      const int64_t f = fid(g, A, i, j, k);
      float u0;
      const float res = sweep_u_face<A>(g, u, ls, mark, sf, i, j, k, f, u0);
      changed = __float_as_uint(res) != __float_as_uint(u0);
      out[f] = res;
    }
    const uint8_t fl = tile_flags(changed);
    if (lane == q) myflag = fl;
  }
  if (cur && lane < per_wave && t0 + lane < ntiles) cur[t0 + lane] = myflag;
}

// setWaterBoundaryU_k spray_k.cu:956-995 (in place).  The ceiling rule reads
// uz(k-1) after that face's own rule.
__device__ __forceinline__ bool isb(const Grid& g, const int8_t* mark, int i, int j, int k) {
  return mark[cid(g, i, j, k)] == BOUNDARY;
}
template <int A>
__global__ void k_bc(Grid g, float* __restrict__ u, const int8_t* __restrict__ mark, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int hh = g.ny + (A == 1), d = g.nz + (A == 2);
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  if (A == 0) {
    if (i <= 1 || i >= (g.nx + 1) - 2) u[f] = 0.0f;
    else if (i > 1 && i < g.nx - 1 && (isb(g, mark, i, j, k) != isb(g, mark, i - 1, j, k))) u[f] = 0.0f;
  } else if (A == 1) {
    if (j <= 1 || j >= (g.ny + 1) - 2) u[f] = 0.0f;
    else if (j > 0 && j < g.ny && (isb(g, mark, i, j, k) != isb(g, mark, i, j - 1, k))) u[f] = 0.0f;
  } else {
    const int zn = g.nz + 1;
    if (k <= 1 || k >= zn - 1) u[f] = 0.0f;
    else if (k == zn - 2) {
      // post-rule value of uz(k-1): k-1 >= 2 is neither wall nor ceiling here
      int kb = k - 1;
      float b = u[f - 1];
      if (kb <= 1) b = 0.0f;
      else if (kb > 0 && kb < g.nz && (isb(g, mark, i, j, kb) != isb(g, mark, i, j, kb - 1))) b = 0.0f;
      u[f] = (b < 0) ? b : 0.0f;
    } else if (k > 0 && k < g.nz && (isb(g, mark, i, j, k) != isb(g, mark, i, j, k - 1))) u[f] = 0.0f;
  }
}

// ---------------- divergence (a11) and gradient (a13) ----------------
__device__ __forceinline__ int8_t mark_at(const Grid& g, const int8_t* mark, int i, int j, int k) {
  return valid(g, i, j, k) ? mark[cid(g, i, j, k)] : BOUNDARY;
}
// cptdivergence_bubble2, spray_k.cu:3138-3263, in arithmetic R.
template <class R>
__global__ void k_divergence(Grid g, const float* __restrict__ wux, const float* __restrict__ wuy,
                             const float* __restrict__ wuz, const float* __restrict__ aux,
                             const float* __restrict__ auy, const float* __restrict__ auz,
                             const int8_t* __restrict__ mark, const float* __restrict__ ls, int solid_c,
                             R* __restrict__ div) {
  CELL_LOOP_HEAD
  R dv = 0;
  const int8_t m = mark[idx];
  if (m == FLUID || m == AIR) {
    const float* wu[3] = {wux, wuy, wuz};
    const float* au[3] = {aux, auy, auz};
    R uu[6];
    int q = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int s = 1; s >= 0; --s) {
        int ni = i + (a == 0 ? (s ? 1 : -1) : 0), nj = j + (a == 1 ? (s ? 1 : -1) : 0), nk = k + (a == 2 ? (s ? 1 : -1) : 0);
        int64_t fx = fid(g, a, i + (a == 0 && s), j + (a == 1 && s), k + (a == 2 && s));
        int8_t mn = mark_at(g, mark, ni, nj, nk);
        R v;
        if (solid_c && mn == SOLID) v = R(0);
        else if (m == FLUID && mn != AIR) v = (R)wu[a][fx];
        else if (m == AIR && mn != FLUID) v = (R)au[a][fx];
        else {
          R lc = (R)ls[idx], ln = (R)ls[cid(g, ni, nj, nk)];
          R th = (R(0.0f) - lc) / (ln - lc);
          if (m == FLUID) v = th * (R)wu[a][fx] + (R(1) - th) * (R)au[a][fx];
          else v = th * (R)au[a][fx] + (R(1) - th) * (R)wu[a][fx];
        }
        uu[q++] = v;
      }
    dv = (uu[0] - uu[1] + uu[2] - uu[3] + uu[4] - uu[5]) / (R)g.h;
  }
  div[idx] = dv;
}

// subGradPress spray_k.cu:198-224 (unit form, in R) or the 1/rho form (double).
template <int A, class R>
__global__ void k_subgrad(Grid g, float* __restrict__ u, const R* __restrict__ p, const int8_t* __restrict__ mark,
                          int vardens, double inv_rho, int solid_c, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int hh = g.ny + (A == 1), d = g.nz + (A == 2);
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const int n[3] = {i, j, k};
  const int lim = (A == 0) ? g.nx : (A == 1) ? g.ny : g.nz;
  if (!(n[A] > 0 && n[A] < lim)) return;
  const int bi = i - (A == 0), bj = j - (A == 1), bk = k - (A == 2);
  const int64_t c0 = cid(g, i, j, k), c1 = cid(g, bi, bj, bk);
  if (solid_c && (mark[c0] == SOLID || mark[c1] == SOLID)) {
    u[f] = 0.0f;
    return;
  }
  if (!vardens) u[f] = (float)((R)u[f] - (p[c0] - p[c1]) / (R)g.h);
  else u[f] = (float)((double)u[f] - inv_rho * (((double)p[c0] - (double)p[c1]) / (double)g.h));
}

// ---------------- heat grid stages (a16) ----------------
// Heat P2G (mapHeatp2g_hash + sumHeat, spray_k.cu:4619-4653) as a tile
// scatter like the liquid P2G: one thread per cell of a 4x8x8 tile, 27
// phases in the gather's order adding each particle to the cell c - o of the
// tile's region (the tile grown by one cell), per-tile sums combined per cell
// in tile order by k_heat_collect.
constexpr int HR_X = ST_X + 2, HR_Y = ST_Y + 2, HR_Z = ST_Z + 2;
constexpr int HR = HR_X * HR_Y * HR_Z;  // 600 cells

// NS particles of cell (ci, cj, ck) in registers, and the phases that add
// them to cell c - o in phase o (the liquid P2G's pass structure).
template <int NS>
__device__ __forceinline__ void heat_load(const f3* __restrict__ pos, const float* __restrict__ temp, const Grid& g,
                                          float scale, uint32_t base, uint32_t cnt, float (&X)[NS], float (&Y)[NS],
                                          float (&Z)[NS], float (&TT)[NS], float (&ON)[NS]) {
  f3 xr[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const uint32_t p = (cnt > 0) ? base + min((uint32_t)q, cnt - 1) : 0u;
    xr[q] = pos[p];
    TT[q] = temp[p];
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of their uses
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const bool on = (uint32_t)q < cnt;
    X[q] = on ? (xr[q].x - g.gx0) * scale : 0.0f;
    Y[q] = on ? (xr[q].y - g.gy0) * scale : 0.0f;
    Z[q] = on ? (xr[q].z - g.gz0) * scale : 0.0f;
    TT[q] = on ? TT[q] : 0.0f;
    ON[q] = on ? 1.0f : 0.0f;
  }
}

template <int NS>
__device__ __forceinline__ void heat_phases(int x0, int y0, int z0, int ci, int cj, int ck, bool act,
                                            const float (&X)[NS], const float (&Y)[NS], const float (&Z)[NS],
                                            const float (&TT)[NS], const float (&ON)[NS], float* sW, float* sH) {
  const float RE = 1.4f, RE2 = RE * RE;
  for (int ph = 0; ph < 27; ++ph) {
    const int f0 = ci - (ph / 9 - 1), f1 = cj - ((ph / 3) % 3 - 1), f2 = ck - (ph % 3 - 1);
    if (act) {
      const int li = ((f0 - (x0 - 1)) * HR_Y + (f1 - (y0 - 1))) * HR_Z + (f2 - (z0 - 1));
      float W = sW[li], Hs = sH[li];
      const float gx = (float)f0 + 0.5f, gy = (float)f1 + 0.5f, gz = (float)f2 + 0.5f;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const float dx = X[q] - gx, dy = Y[q] - gy, dz = Z[q] - gz;
        const float w = ON[q] * fmaxf(div_nr(RE2, fmaxf(dx * dx + dy * dy + dz * dz, 0.0001f)) - 1.0f, 0.0f);
        W += w;
        Hs += w * TT[q];
      }
      sW[li] = W;
      sH[li] = Hs;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(ST_CELLS)
    k_heat_tile(Grid g, const f3* __restrict__ pos, const float* __restrict__ temp,
                const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                const uint8_t* __restrict__ cmask, float* __restrict__ parts, int ntx, int nty, int ntz) {
  __shared__ float sW[HR];
  __shared__ float sH[HR];
  __shared__ uint32_t sS[ST_CELLS], sN[ST_CELLS], sList[ST_CELLS];
  __shared__ int sWc[ST_CELLS / 64];
  const int tl = (int)xcd_block();  // XCD-contiguous tiles
  int t = tl;
  const int tz = t % ntz;
  t /= ntz;
  const int ty = t % nty;
  const int tx = t / nty;
  const int x0 = tx * ST_X, y0 = ty * ST_Y, z0 = tz * ST_Z;
  for (int q = threadIdx.x; q < HR; q += ST_CELLS) {
    sW[q] = 0.0f;
    sH[q] = 0.0f;
  }
  const int b = threadIdx.x;
  const int ci = x0 + b / (ST_Y * ST_Z), cj = y0 + (b / ST_Z) % ST_Y, ck = z0 + b % ST_Z;
  uint32_t s = 0, e = 0;
  if (valid(g, ci, cj, ck)) {
    const int64_t c = cid(g, ci, cj, ck);
    if (cmask[c]) {
      s = cstart[c];
      e = cend[c];
    }
  }
  const int any = __syncthreads_or(e > s);
  if (any) {
    const float scale = 1.0f / g.h;
    const uint32_t n = e - s;
    sS[b] = s;
    sN[b] = n;
    {
      float X[SCH], Y[SCH], Z[SCH], TT[SCH], ON[SCH];
      heat_load<SCH>(pos, temp, g, scale, s, min(n, (uint32_t)SCH), X, Y, Z, TT, ON);
      heat_phases<SCH>(x0, y0, z0, ci, cj, ck, n > 0, X, Y, Z, TT, ON, sW, sH);
    }
    overflow_passes(n, sList, sWc, [&](bool act, uint32_t cb, uint32_t thr) {
      const int oi = x0 + (int)cb / (ST_Y * ST_Z), oj = y0 + ((int)cb / ST_Z) % ST_Y, ok = z0 + (int)cb % ST_Z;
      const uint32_t os = act ? sS[cb] + thr : 0u;
      const uint32_t oc = act ? min(sN[cb] - thr, (uint32_t)SCH2) : 0u;
      float X[SCH2], Y[SCH2], Z[SCH2], TT[SCH2], ON[SCH2];
      if (act) heat_load<SCH2>(pos, temp, g, scale, os, oc, X, Y, Z, TT, ON);
      heat_phases<SCH2>(x0, y0, z0, oi, oj, ok, act, X, Y, Z, TT, ON, sW, sH);
    });
  } else {
    __syncthreads();
  }
  float* dst = parts + (int64_t)tl * (2 * HR);
  for (int q = threadIdx.x; q < HR; q += ST_CELLS) {
    dst[q] = sW[q];
    dst[HR + q] = sH[q];
  }
}

// T of a cell from the regions of the (at most 8) tiles covering it, in tile
// order; cells without particle weight keep the default temperature.
__global__ void k_heat_collect(Grid g, const float* __restrict__ parts, int ntx, int nty, int ntz, float deft,
                               float* __restrict__ tp) {
  CELL_LOOP_HEAD
  const int cv[3] = {i, j, k};
  const int nt[3] = {ntx, nty, ntz};
  int qlo[3], qhi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int T = st_t(d);  // region [qT - 1, qT + T]
    qlo[d] = (cv[d] - T < 0) ? 0 : (cv[d] - 1) / T;
    const int hi = (cv[d] + 1) / T;
    qhi[d] = hi > nt[d] - 1 ? nt[d] - 1 : hi;
  }
  float W = 0.0f, Hs = 0.0f;
  for (int qx = qlo[0]; qx <= qhi[0]; ++qx)
    for (int qy = qlo[1]; qy <= qhi[1]; ++qy)
      for (int qz = qlo[2]; qz <= qhi[2]; ++qz) {
        const int li = ((i - (qx * ST_X - 1)) * HR_Y + (j - (qy * ST_Y - 1))) * HR_Z + (k - (qz * ST_Z - 1));
        const float* src = parts + ((int64_t)(qx * nty + qy) * ntz + qz) * (2 * HR);
        W += src[li];
        Hs += src[HR + li];
      }
  tp[idx] = (W > 0) ? (Hs / W) : deft;
}

// addHeatAtBottom spray_k.cu:4492-4504, then save T and form b (compb_heat
// 4507-4523) for the implicit solve.
__global__ void k_heat_pre(Grid g, float* __restrict__ tp, float* __restrict__ tsave, float* __restrict__ b,
                           const int8_t* __restrict__ mark, int add_bottom, float inc, float clampv, float a0, float a1,
                           float a2, float a3, float a4, float a5, float dt) {
  CELL_LOOP_HEAD
  float t = tp[idx];
  if (add_bottom) {
    if (i >= 1 && i < g.nx - 1 && j >= 1 && j < g.ny - 1 && k <= 3) t += inc;
    t = fminf(clampv, t);
  }
  tp[idx] = t;
  tsave[idx] = t;
  const float al[6] = {a0, a1, a2, a3, a4, a5};
  int8_t m = mark[idx];
  b[idx] = (m == BOUNDARY) ? 0.0f : (float)((double)t * ((double)g.h * (double)g.h) / (double)al[m] / (double)dt);
}

// One Jacobi sweep of the heat operator (computeAx_heat, spray_k.cu:4526-4567).
// Per cell, once per Jacobi solve: its type (7 = BOUNDARY) and the 6-bit
// mask of neighbours that are not Neumann, in the order +x, +y, +z, -x, -y, -z.
__global__ void k_heat_code(Grid g, const int8_t* __restrict__ mark, uint16_t* __restrict__ code, int vac_neumann) {
  CELL_LOOP_HEAD
  const int8_t m = mark[idx];
  if (m == BOUNDARY) {
    code[idx] = 7;
    return;
  }
  const int off[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  unsigned mask = 0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int8_t mn = mark_at(g, mark, i + off[q][0], j + off[q][1], k + off[q][2]);
    if (!(mn == BOUNDARY || (vac_neumann && mn == VACUUM))) mask |= 1u << q;
  }
  code[idx] = (uint16_t)((unsigned)m | (mask << 3));
}

// One Jacobi pass: the diagonal h^2/(alpha_m dt) + 6, less one per Neumann
// neighbour, from a per-block table of the 6 x 7 (type, Neumann count)
// values (each subtraction of 1 is exact, so the table equals the
// one-by-one form); x of the open neighbours added to b in face order.
__global__ void k_heat_jacobi(Grid g, const float* __restrict__ x, float* __restrict__ xn, const float* __restrict__ b,
                              const uint16_t* __restrict__ code, float a0, float a1, float a2, float a3, float a4,
                              float a5, float dt) {
  __shared__ double dgt[6][7];
  if (threadIdx.x < 42) {
    const float al[6] = {a0, a1, a2, a3, a4, a5};
    const int m = threadIdx.x / 7, nn = threadIdx.x % 7;
    double d = (double)g.h * (double)g.h / (double)al[m] / (double)dt + 6.0;
    for (int q = 0; q < nn; ++q) d -= 1.0;
    dgt[m][nn] = d;
  }
  __syncthreads();
  CELL_LOOP_HEAD
  const unsigned cd = code[idx];
  if ((cd & 7u) == 7u) {
    xn[idx] = 0.0f;
    return;
  }
  const unsigned mask = cd >> 3;
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const int64_t st[6] = {sx, sy, 1, -sx, -sy, -1};
  // the six neighbour loads issued together (a closed face reads the cell)
  float xv[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) xv[q] = x[(mask & (1u << q)) ? idx + st[q] : idx];
  double s = (double)b[idx];
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if (mask & (1u << q)) s += (double)xv[q];
  xn[idx] = (float)(s / dgt[cd & 7u][6 - __popc(mask)]);
}

__global__ void k_heat_start(int64_t G, const float* __restrict__ tsave, const int8_t* __restrict__ mark,
                             float* __restrict__ x) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < G) x[idx] = (mark[idx] == BOUNDARY) ? 0.0f : tsave[idx];
}

// setBoundaryHeat spray_k.cu:4585-4605 (reads the pass's input) ; compTpChange 4607-4617
__global__ void k_heat_post(Grid g, const float* __restrict__ tin, float* __restrict__ tp,
                            const float* __restrict__ tsave, float* __restrict__ tch, const int8_t* __restrict__ mark,
                            int extend) {
  CELL_LOOP_HEAD
  float t = tin[idx];
  if (extend) {
    if (i == g.nx - 1) t = tin[cid(g, i - 1, j, k)];
    else if (i == 0) t = tin[cid(g, i + 1, j, k)];
    else if (j == g.ny - 1) t = tin[cid(g, i, j - 1, k)];
    else if (j == 0) t = tin[cid(g, i, j + 1, k)];
    else if (k == g.nz - 1) t = tin[cid(g, i, j, k - 1)];
    else if (k == 0) t = tin[cid(g, i, j, k + 1)];
  }
  tp[idx] = t;
  tch[idx] = (mark[idx] != BOUNDARY) ? t - tsave[idx] : 0.0f;
}

// ---------------- solid distance + velocity (a15 support) ----------------
// initSolidPhi spray_k.cu:724-744
__global__ void k_init_solid_phi(int64_t G, const uint8_t* __restrict__ cmask, float* __restrict__ phi, float big) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < G) phi[idx] = (cmask[idx] & M_SOLID) ? -0.5f : big;
}

// ---------------- marching cubes (a18) ----------------
// A node lattice: nodes (nx, ny, nz), voxels one fewer per axis; node
// (i,j,k) at index (i*ny + j)*nz + k, position (i + off) * h + origin.  The
// level-set source uses cell centres of the N grid (off 0.5); the particle
// sources the (2N+1)^3 nodes of spacing dx/2 (off 0, spray_k.cu:1649).
struct McGrid {
  int nx, ny, nz;
  int64_t N, V;
  float h, off, x0, y0, z0;
};
__device__ __forceinline__ int64_t nid(const McGrid& m, int i, int j, int k) {
  return ((int64_t)i * m.ny + j) * m.nz + k;
}

// Cube edges as (di, dj, dk, slot): the edge starts at node (i+di, j+dj, k+dk)
// and runs along +x (slot 1), +y (slot 0) or +z (slot 2); its id is
// 3 * node + slot, the numbering of GetEdgeID (spray_k.cu:1920-1957).
__constant__ int c_eo[12][4] = {{0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 1}, {0, 0, 0, 0}, {0, 0, 1, 1}, {1, 0, 1, 0},
                                {0, 1, 1, 1}, {0, 0, 1, 0}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
__constant__ signed char c_tri[256][16];
__constant__ unsigned char c_nv[256];

// classifyVoxel, spray_k.cu:1853-1888 (isoValue 0)
__device__ __forceinline__ uint32_t mc_case_of(const McGrid& m, const float* f, int i, int j, int k) {
  const int64_t si = (int64_t)m.ny * m.nz, sj = m.nz;
  const float* b = f + nid(m, i, j, k);
  float v[8] = {b[0], b[si], b[si + sj], b[sj], b[1], b[si + 1], b[si + sj + 1], b[sj + 1]};
  uint32_t ci = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) ci += (uint32_t)(v[q] < 0.0f) << q;
  return ci;
}

// Case index, triangle count and the edges its triangles use
// (markActiveEdge_MC, spray_k.cu:2172-2220).
__global__ void k_mc_classify(McGrid m, const float* __restrict__ f, uint8_t* __restrict__ cases,
                              uint32_t* __restrict__ tcount, uint32_t* __restrict__ edge_used,
                              uint32_t* __restrict__ active) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int nv = 0;
  if (v < m.V) {
    int i, j, k;
    ijk3(v, m.ny - 1, m.nz - 1, i, j, k);
    uint32_t ci = mc_case_of(m, f, i, j, k);
    cases[v] = (uint8_t)ci;
    nv = c_nv[ci];
    tcount[v] = (uint32_t)(nv / 3);
    for (int t = 0; t < nv; ++t) {
      int e = c_tri[ci][t];
      edge_used[3 * nid(m, i + c_eo[e][0], j + c_eo[e][1], k + c_eo[e][2]) + c_eo[e][3]] = 1u;
    }
  }
  // active voxels: one add per block, spread over NSPREAD counters
  const int cnt = __syncthreads_count(nv > 0);
  if (threadIdx.x == 0 && cnt) atomicAdd(active + (blockIdx.x % NSPREAD) * SPREAD_STRIDE, (uint32_t)cnt);
}

// vertexInterp (spray_k.cu:1903-1908) along each used edge, from its start
// node p0 to p1 = p0 + h along the edge (generateTriangles_indices,
// spray_k.cu:2092-2117); vid holds the exclusive scan of the used flags.  The
// reference writes a shared edge's vertex from whichever voxel comes last;
// here every edge is interpolated once, in the + direction.
__global__ void k_mc_verts(McGrid m, const float* __restrict__ f, const uint32_t* __restrict__ used,
                           const uint32_t* __restrict__ vid, float* __restrict__ verts,
                           uint32_t* __restrict__ vedge) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * m.N || !used[e]) return;
  const int64_t n = e / 3;
  const int slot = (int)(e % 3);
  const int a = slot == 1 ? 0 : (slot == 0 ? 1 : 2);
  int i, j, k;
  ijk3(n, m.ny, m.nz, i, j, k);
  const int i2 = i + (a == 0), j2 = j + (a == 1), k2 = k + (a == 2);
  const float f0 = f[n], f1 = f[nid(m, i2, j2, k2)];
  const float t = (0.0f - f0) / (f1 - f0);
  const float h = m.h;
  const float p0x = ((float)i + m.off) * h + m.x0, p0y = ((float)j + m.off) * h + m.y0,
              p0z = ((float)k + m.off) * h + m.z0;
  const float p1x = a == 0 ? p0x + h : p0x, p1y = a == 1 ? p0y + h : p0y, p1z = a == 2 ? p0z + h : p0z;
  const uint32_t o = vid[e];
  vedge[o] = (uint32_t)e;
  verts[3 * (int64_t)o + 0] = p0x + t * (p1x - p0x);
  verts[3 * (int64_t)o + 1] = p0y + t * (p1y - p0y);
  verts[3 * (int64_t)o + 2] = p0z + t * (p1z - p0z);
}

// generateTriangles_indices spray_k.cu:2074-2170 (indices = scanned edge ids)
__global__ void k_mc_tris(McGrid m, const uint8_t* __restrict__ cases, const uint32_t* __restrict__ toff,
                          const uint32_t* __restrict__ vid, uint32_t* __restrict__ tris) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m.V) return;
  uint32_t ci = cases[v];
  int nv = c_nv[ci];
  if (!nv) return;
  int i, j, k;
  ijk3(v, m.ny - 1, m.nz - 1, i, j, k);
  int64_t w = 3 * (int64_t)toff[v];
  for (int t = 0; t < nv; ++t) {
    int e = c_tri[ci][t];
    tris[w + t] = vid[3 * nid(m, i + c_eo[e][0], j + c_eo[e][1], k + c_eo[e][2]) + c_eo[e][3]];
  }
}

// The triangles around one vertex, in the order of the triangle list: the
// vertex lies on edge e, which the (at most) 4 voxels around it share; their
// index order is the list's order, and within a voxel its triangles follow
// the table.  fn(i1, i2, i3) sees every triangle that holds the vertex.
template <class Fn>
__device__ __forceinline__ void mc_vertex_triangles(const McGrid& m, uint32_t e, const uint8_t* __restrict__ cases,
                                                    const uint32_t* __restrict__ toff,
                                                    const uint32_t* __restrict__ tris, uint32_t v, Fn fn) {
  const int64_t n = e / 3;
  const int slot = (int)(e % 3);
  int i, j, k;
  ijk3(n, m.ny, m.nz, i, j, k);
  const int VX = m.nx - 1, VY = m.ny - 1, VZ = m.nz - 1;
  for (int q = 0; q < 4; ++q) {
    const int a = q >> 1, b = q & 1;
    int vi, vj, vk;
    if (slot == 1) { vi = i; vj = j - 1 + a; vk = k - 1 + b; }       // x edge
    else if (slot == 0) { vi = i - 1 + a; vj = j; vk = k - 1 + b; }  // y edge
    else { vi = i - 1 + a; vj = j - 1 + b; vk = k; }                  // z edge
    if (vi < 0 || vi >= VX || vj < 0 || vj >= VY || vk < 0 || vk >= VZ) continue;
    const int64_t vox = ((int64_t)vi * VY + vj) * VZ + vk;
    const int nt = c_nv[cases[vox]] / 3;
    const int64_t t0 = toff[vox];
    for (int t = 0; t < nt; ++t) {
      const uint32_t i1 = tris[3 * (t0 + t)], i2 = tris[3 * (t0 + t) + 1], i3 = tris[3 * (t0 + t) + 2];
      if (i1 == v || i2 == v || i3 == v) fn(i1, i2, i3);
    }
  }
}

// calNormals, spray.cu:981-997: the sum of cross(p2 - p1, p3 - p1) over the
// vertex's triangles (calnormal_k, spray_k.cu:2223-2250, adds them with
// atomics; here in the list's order), normalised when nonzero
// (normalizeTriangleNor_k, 2252-2261).
__global__ void k_mc_normals(McGrid m, const uint32_t* __restrict__ vedge, const uint8_t* __restrict__ cases,
                             const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tris,
                             const float* __restrict__ verts, float* __restrict__ nrm, int64_t nv) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  mc_vertex_triangles(m, vedge[v], cases, toff, tris, (uint32_t)v, [&](uint32_t i1, uint32_t i2, uint32_t i3) {
    const float* a = verts + 3 * (int64_t)i1;
    const float* b = verts + 3 * (int64_t)i2;
    const float* c = verts + 3 * (int64_t)i3;
    const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const float wx = c[0] - a[0], wy = c[1] - a[1], wz = c[2] - a[2];
    nx = nx + (uy * wz - uz * wy);
    ny = ny + (uz * wx - ux * wz);
    nz = nz + (ux * wy - uy * wx);
  });
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  if (len > 0.0f) {
    const float inv = 1.0f / len;
    nx = nx * inv;
    ny = ny * inv;
    nz = nz * inv;
  }
  nrm[3 * v] = nx;
  nrm[3 * v + 1] = ny;
  nrm[3 * v + 2] = nz;
}

// smoothMesh, spray.cu:1271-1288, one iteration: each vertex gathers the
// edge vectors to the other two corners of its triangles (in the order of
// smooth_computedisplacement, spray_k.cu:2408-2427) and moves by
// lambda * sum / count (smooth_addDisplacement, 2429-2440).  The
// reference's second, mu = -0.53 pass finds the sums already cleared by the
// first and does nothing; so here.  Positions of the pass's input are read.
__global__ void k_mc_smooth(McGrid m, const uint32_t* __restrict__ vedge, const uint8_t* __restrict__ cases,
                            const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tris,
                            const float* __restrict__ vin, float* __restrict__ vout, int64_t nv, float lambda) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  int w = 0;
  const float* pv = vin + 3 * v;
  mc_vertex_triangles(m, vedge[v], cases, toff, tris, (uint32_t)v, [&](uint32_t i1, uint32_t i2, uint32_t i3) {
    const uint32_t o1 = (i1 == (uint32_t)v) ? i2 : i1;
    const uint32_t o2 = (i3 == (uint32_t)v) ? i2 : i3;
    const float* a = vin + 3 * (int64_t)o1;
    const float* b = vin + 3 * (int64_t)o2;
    dx = dx + (a[0] - pv[0]);
    dy = dy + (a[1] - pv[1]);
    dz = dz + (a[2] - pv[2]);
    dx = dx + (b[0] - pv[0]);
    dy = dy + (b[1] - pv[1]);
    dz = dz + (b[2] - pv[2]);
    w += 2;
  });
  float x = pv[0], y = pv[1], z = pv[2];
  if (w > 0) {
    const float fw = (float)w;
    x = x + lambda * dx / fw;
    y = y + lambda * dy / fw;
    z = z + lambda * dz / fw;
  }
  vout[3 * v] = x;
  vout[3 * v + 1] = y;
  vout[3 * v + 2] = z;
}

// Node field of a particle source, one thread per node of the (2N+1)^3
// lattice, reading the 4x4x4 cells of the 2N lists around it (di, dj, dk in
// [-2, 1]):
//  MPF_MC_LIQUID_GAS  genWaterDensfield_liquidAndGas + sumcelldens_liquidAndGas
//                     (spray_k.cu:1595-1686): FLUID/AIR/AIRSOLO within R = 2h,
//                     w = (R^2 - d^2)^3, phi = h/2 - |x - centre| (-h/2 when
//                     no particle), lattice faces -1000;
//  MPF_MC_GAS         genWaterDensfield_Gas + sumcelldens_Gas (1528-1593): AIR
//                     and, when airsolo, AIRSOLO; r = 0.8h;
//  MPF_MC_SOLID       genWaterDensfield_GY + sumcelldens3 (1688-1772): SOLID
//                     within R = h, w = max((1 - (d/4h)^2)^3, 0), phi =
//                     |x - centre| - 0.75h f(Ev) with the centre measured
//                     from the origin (the reference's centertmp is zero),
//                     lattice faces min(phi, -10).
// Per cell the weighted positions are summed from zero and added to the
// centre, the weights accumulate directly (the reference's order).
template <int SRC>
__global__ void k_mc_density(McGrid m, Grid g2, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                             const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
                             float* __restrict__ out, int airsolo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m.N) return;
  int i, j, k;
  ijk3(idx, m.ny, m.nz, i, j, k);
  const float h = m.h;
  const float px = ((float)i + m.off) * h + m.x0, py = ((float)j + m.off) * h + m.y0,
              pz = ((float)k + m.off) * h + m.z0;
  const float R = (SRC == MPF_MC_SOLID) ? h : h * 2.0f;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f, wsum = 0.0f;
  for (int di = -2; di <= 1; ++di)
    for (int dj = -2; dj <= 1; ++dj)
      for (int dk = -2; dk <= 1; ++dk) {
        const int a = i + di, b = j + dj, q = k + dk;
        if (a < 0 || a > g2.nx - 1 || b < 0 || b > g2.ny - 1 || q < 0 || q > g2.nz - 1) continue;
        const int64_t cell = ((int64_t)a * g2.ny + b) * g2.nz + q;
        const uint32_t s = cs[cell];
        if (s == MPF_CELL_UNDEF) continue;
        const uint32_t e = ce[cell];
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
        for (uint32_t p = s; p < e; ++p) {
          const int8_t fl = flag[p];
          bool take;
          if (SRC == MPF_MC_LIQUID_GAS) take = fl == AIR || fl == AIRSOLO || fl == FLUID;
          else if (SRC == MPF_MC_GAS) take = fl == AIR || (fl == AIRSOLO && airsolo);
          else take = fl == SOLID;
          if (!take) continue;
          const f3 x = pos[p];
          const float dx = x.x - px, dy = x.y - py, dz = x.z - pz;
          const float dis = sqrtf(dx * dx + dy * dy + dz * dz);
          if (dis < R) {
            float w;
            if (SRC == MPF_MC_SOLID) {
              w = dis / (4.0f * h);
              w = 1.0f - w * w;
              w = fmaxf(w * w * w, 0.0f);
            } else {
              w = R * R - dis * dis;
              w = w * w * w;
            }
            rx = rx + x.x * w;
            ry = ry + x.y * w;
            rz = rz + x.z * w;
            wsum = wsum + w;
          }
        }
        cx = cx + rx;
        cy = cy + ry;
        cz = cz + rz;
      }
  const float r = (SRC == MPF_MC_LIQUID_GAS) ? 0.5f * h : (SRC == MPF_MC_GAS ? 0.8f * h : 0.75f * h);
  float phi;
  if (wsum > 0.0f) {
    cx = cx / wsum;
    cy = cy / wsum;
    cz = cz / wsum;
    const float ex = px - cx, ey = py - cy, ez = pz - cz;
    const float len = sqrtf(ex * ex + ey * ey + ez * ez);
    if (SRC == MPF_MC_SOLID) {
      const float thigh = 0.51f, tlow = 0.49f;
      const float Ev = fmaxf(cx, fmaxf(cy, cz)) / (4.0f * h);
      const float gamma = (thigh - Ev) / (thigh - tlow);
      const float fr = (Ev < tlow) ? 1.0f : gamma * gamma * gamma - 3.0f * gamma * gamma + 3.0f * gamma;
      phi = len - r * fr;
    } else {
      phi = r - len;
    }
  } else {
    phi = -r;
  }
  const bool face = i == 0 || j == 0 || k == 0 || i == m.nx - 1 || j == m.ny - 1 || k == m.nz - 1;
  if (face) phi = (SRC == MPF_MC_SOLID) ? fminf(phi, -10.0f) : -1000.0f;
  out[idx] = phi;
}

}  // namespace

// ================= host stage drivers =================
// n two-buffer passes, ping-ponging between phi and the scratch field (one
// copy at the end when n is odd).
// n passes ending in phi; the initial field is in c.ctmp when in_tmp (the
// preparing kernel wrote it there, so an odd n needs no final copy)
static void sweep_n(Ctx& c, float* phi, int typeflag, int n, bool in_tmp = false) {
  float* src = in_tmp ? c.ctmp : phi;
  float* dst = in_tmp ? phi : c.ctmp;
  const uint32_t ntj = (uint32_t)((c.g.ny + SW_J - 1) / SW_J), ntk = (uint32_t)((c.g.nz + SW_K - 1) / SW_K);
  const int64_t nt = sweep_tiles(c.g);
  const int64_t nwaves = (nt + SW_L - 1) / SW_L;
  for (int it = 0; it < n; ++it) {
    const uint8_t* prev = it == 0 ? nullptr : c.sflag + ((it - 1) & 1) * nt;
    launch(c, "sweep", dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, k_sweep, c.g, (const float*)src, dst,
           (const int8_t*)c.mark, typeflag, prev, c.sflag + (it & 1) * nt, ntj, ntk, (uint32_t)nt);
    float* t = src;
    src = dst;
    dst = t;
  }
  if (src != phi) hipMemcpyAsync(phi, src, (size_t)c.g.G * 4, hipMemcpyDeviceToDevice, c.stream);
}

static void redistance(Ctx& c, float* ls) {
  const int64_t G = c.g.G;
  launch(c, "ls_prepare", dim3(nblocks(G, 256)), dim3(256), 0, k_ls_prepare, G, (const float*)ls, c.ctmp, c.mark,
         c.g.h, 0);
  sweep_n(c, ls, FLUID, 5, true);
  launch(c, "ls_back", dim3(nblocks(G, 256)), dim3(256), 0, k_ls_back, G, ls, c.g.h, 0);
  launch(c, "ls_prepare", dim3(nblocks(G, 256)), dim3(256), 0, k_ls_prepare, G, (const float*)ls, c.ctmp, c.mark,
         c.g.h, 1);
  sweep_n(c, ls, FLUID, 5, true);
  launch(c, "ls_back", dim3(nblocks(G, 256)), dim3(256), 0, k_ls_back, G, ls, c.g.h, 1);
}

static void upload_ring_tables() {
  static bool done = false;
  if (done) return;
  struct O { int d[3]; float lb; };
  std::vector<O> offs;
  for (int di = -2; di <= 2; ++di)
    for (int dj = -2; dj <= 2; ++dj)
      for (int dk = -2; dk <= 2; ++dk) {
        int d[3] = {di, dj, dk};
        float lb = 0;
        for (int a = 0; a < 3; ++a) {
          float m = fabsf((float)d[a]) - 0.5f;
          if (m > 0) lb += m * m;
        }
        if (lb > 6.25f) continue;  // corners of the window: never below 2.5h
        offs.push_back(O{{di, dj, dk}, lb});
      }
  std::stable_sort(offs.begin(), offs.end(), [](const O& x, const O& y) { return x.lb < y.lb; });
  int tab[120] = {0};
  int start[10];
  float lbs[9];
  int nr = 0;
  for (size_t q = 0; q < offs.size(); ++q) {
    if (q == 0 || offs[q].lb != offs[q - 1].lb) {
      start[nr] = (int)q;
      lbs[nr] = offs[q].lb;
      ++nr;
    }
    tab[q] = (offs[q].d[0] + 4) | ((offs[q].d[1] + 4) << 8) | ((offs[q].d[2] + 4) << 16);
  }
  start[nr] = (int)offs.size();
  hipMemcpyToSymbol(HIP_SYMBOL(c_ring_off), tab, sizeof(tab));
  hipMemcpyToSymbol(HIP_SYMBOL(c_ring_start), start, sizeof(start));
  hipMemcpyToSymbol(HIP_SYMBOL(c_ring_lb), lbs, sizeof(lbs));
  done = true;
}

// Brick summary of the occupancy masks (after every change of the masks).
void run_bricks(Ctx& c) {
  const int bx = (c.g.nx + 3) / 4, by = (c.g.ny + 3) / 4, bz = (c.g.nz + 3) / 4;
  const int64_t NB = nbricks(c.g);
  launch(c, "bricks", dim3(nblocks(NB, 256)), dim3(256), 0, k_brick_or, c.g, (const uint8_t*)c.cmask, c.bmask, bx, by,
         bz);
  launch(c, "bricks", dim3(nblocks(NB, 256)), dim3(256), 0, k_brick_dilate, (const uint8_t*)c.bmask, c.bmask + NB, bx,
         by, bz);
}

void run_levelset(Ctx& c) {
  upload_ring_tables();
  const int64_t G = c.g.G;
  const int cur = c.cur;
  const int64_t NB = nbricks(c.g);
  launch(c, "genlevelset", dim3(nblocks(G, 256)), dim3(256), 0, k_genlevelset, c.g, (const f3*)c.pos[cur],
         (const int8_t*)c.flag[cur], (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask,
         (const uint8_t*)(c.bmask + NB), c.lsfluid, c.lsair, c.mark);
  redistance(c, c.lsfluid);
  redistance(c, c.lsair);
  launch(c, "merge_mark", dim3(nblocks(G, 256)), dim3(256), 0, k_merge, c.g, (const float*)c.lsfluid,
         (const float*)c.lsair, c.lsmerge, c.mark);
  launch(c, "phigrad", dim3(nblocks(G, 256)), dim3(256), 0, k_phigrad, c.g, (const float*)c.lsmerge, c.phig[0],
         c.phig[1], c.phig[2]);
  if (c.cfg.flags_from_counts)
    launch(c, "mark_counts", dim3(nblocks(G, 256)), dim3(256), 0, k_mark_counts, c.g, (const uint32_t*)c.cstart,
           (const uint32_t*)c.cend, (const int8_t*)c.flag[cur], c.mark);
  if (c.cfg.solid_constraint)
    launch(c, "mark_solid", dim3(nblocks(G, 256)), dim3(256), 0, k_mark_solid, c.g, (const uint32_t*)c.cstart,
           (const uint32_t*)c.cend, (const int8_t*)c.flag[cur], (const uint8_t*)c.cmask, c.mark);
}

template <bool AIRPH>
static void p2g_phase(Ctx& c, float* const* out, float* const* out2) {
  const int cur = c.cur;
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  const f3* pos = c.pos[cur];
  const f3* vel = c.vel[cur];
  const float* mass = c.mass[cur];
  const int8_t* flag = c.flag[cur];
  const char* nm = AIRPH ? "p2g_air" : "p2g_liquid";
  launch(c, nm, dim3(nblocks(fc[0], 256)), dim3(256), 0, k_p2g<0, AIRPH>, c.g, pos, vel, mass, flag,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask,
         (const uint8_t*)(c.bmask + nbricks(c.g)), out[0], out2 ? out2[0] : (float*)nullptr, fc[0]);
  launch(c, nm, dim3(nblocks(fc[1], 256)), dim3(256), 0, k_p2g<1, AIRPH>, c.g, pos, vel, mass, flag,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask,
         (const uint8_t*)(c.bmask + nbricks(c.g)), out[1], out2 ? out2[1] : (float*)nullptr, fc[1]);
  launch(c, nm, dim3(nblocks(fc[2], 256)), dim3(256), 0, k_p2g<2, AIRPH>, c.g, pos, vel, mass, flag,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask,
         (const uint8_t*)(c.bmask + nbricks(c.g)), out[2], out2 ? out2[2] : (float*)nullptr, fc[2]);
}

static void p2g_liquid(Ctx& c, float* const* out, float* const* out2 = nullptr, const float* band = nullptr,
                       float big = 0.0f) {
  const int cur = c.cur;
  const int ntx = (c.g.nx + ST_X - 1) / ST_X, nty = (c.g.ny + ST_Y - 1) / ST_Y, ntz = (c.g.nz + ST_Z - 1) / ST_Z;
  launch(c, band ? "p2g_solid_band" : "p2g_liquid", dim3((unsigned)(ntx * nty * ntz)), dim3(ST_CELLS), 0,
         k_p2g_tile, c.g, (const f3*)c.pos[cur], (const f3*)c.vel[cur], (const float*)c.mass[cur],
         (const int8_t*)c.flag[cur], (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask,
         c.p2g_parts, ntx, nty, ntz, band, big, (const uint8_t*)(c.bmask + nbricks(c.g)));
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  const char* nm = band ? "p2g_solid_collect" : "p2g_collect";
  launch(c, nm, dim3(nblocks(fc[0], 256)), dim3(256), 0, k_p2g_collect<0>, c.g, (const float*)c.p2g_parts, ntx, nty,
         ntz, out[0], out2 ? out2[0] : (float*)nullptr, fc[0], band, big);
  launch(c, nm, dim3(nblocks(fc[1], 256)), dim3(256), 0, k_p2g_collect<1>, c.g, (const float*)c.p2g_parts, ntx, nty,
         ntz, out[1], out2 ? out2[1] : (float*)nullptr, fc[1], band, big);
  launch(c, nm, dim3(nblocks(fc[2], 256)), dim3(256), 0, k_p2g_collect<2>, c.g, (const float*)c.p2g_parts, ntx, nty,
         ntz, out[2], out2 ? out2[2] : (float*)nullptr, fc[2], band, big);
}

int64_t p2g_parts_floats(const Grid& g) {
  const int64_t nt = (int64_t)((g.nx + ST_X - 1) / ST_X) * ((g.ny + ST_Y - 1) / ST_Y) * ((g.nz + ST_Z - 1) / ST_Z);
  return nt * 2 * ST_RTOT;
}

void run_p2g(Ctx& c) {
  p2g_liquid(c, c.wu, c.wuo);
  p2g_phase<true>(c, c.au, c.auo);
}

void run_solid_phi_and_velocity(Ctx& c) {
  const int64_t G = c.g.G;
  launch(c, "init_solid_phi", dim3(nblocks(G, 256)), dim3(256), 0, k_init_solid_phi, G, (const uint8_t*)c.cmask,
         c.ctmp, (float)(3 * c.g.nx));
  sweep_n(c, c.phisolid, -1, 3, true);
  // solid velocity only where the collision response can read it
  p2g_liquid(c, c.su, nullptr, c.phisolid, (float)(3 * c.g.nx));
}

static void sweep_phi(Ctx& c, float* phi, int8_t type) {
  const int64_t G = c.g.G;
  launch(c, "initphi", dim3(nblocks(G, 256)), dim3(256), 0, k_initphi, G, (const int8_t*)c.mark, c.ctmp, type,
         (float)(c.g.nx * 3));
  sweep_n(c, phi, type, 3, true);
}

// Two passes: u -> scratch, scratch -> u.
template <int A>
static void sweep_u_pass(Ctx& c, const float* src, float* dst, const float* phi, int8_t sf, const uint8_t* prev,
                         uint8_t* cur) {
  const int64_t nfy = c.g.ny + (A == 1), nfz = c.g.nz + (A == 2);
  const uint32_t ntj = (uint32_t)((nfy + SW_J - 1) / SW_J), ntk = (uint32_t)((nfz + SW_K - 1) / SW_K);
  const int64_t nt = face_tiles(c.g, A);
  // a pass that computes every tile gives each wave SW_U1 tiles (more waves
  // in flight for its two rounds of loads); a pass that skips reads the
  // flags of SW_L tiles per wave
  const int per_wave = prev ? SW_L : SW_U1;
  const int64_t nwaves = (nt + per_wave - 1) / per_wave;
  launch(c, "sweep_u", dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, k_sweep_u<A>, c.g, src, dst, phi,
         (const int8_t*)c.mark, sf, prev, cur, ntj, ntk, (uint32_t)nt, per_wave);
}
static void sweep_u(Ctx& c, float* const* u, const float* phi, int8_t sf) {
  // the first pass records per-tile changes (one flag set per axis), the
  // second skips the tiles they leave unchanged
  uint8_t* fl[3] = {c.sflag + 2 * sweep_tiles(c.g), nullptr, nullptr};
  fl[1] = fl[0] + face_tiles(c.g, 0);
  fl[2] = fl[1] + face_tiles(c.g, 1);
  for (int it = 0; it < 2; ++it) {
    float* const* src = it == 0 ? u : c.ftmp;
    float* const* dst = it == 0 ? c.ftmp : u;
    const bool first = it == 0;
    sweep_u_pass<0>(c, src[0], dst[0], phi, sf, first ? nullptr : fl[0], first ? fl[0] : nullptr);
    sweep_u_pass<1>(c, src[1], dst[1], phi, sf, first ? nullptr : fl[1], first ? fl[1] : nullptr);
    sweep_u_pass<2>(c, src[2], dst[2], phi, sf, first ? nullptr : fl[2], first ? fl[2] : nullptr);
  }
}

static void bc(Ctx& c, float* const* u) {
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  launch(c, "bc", dim3(nblocks(fc[0], 256)), dim3(256), 0, k_bc<0>, c.g, u[0], (const int8_t*)c.mark, fc[0]);
  launch(c, "bc", dim3(nblocks(fc[1], 256)), dim3(256), 0, k_bc<1>, c.g, u[1], (const int8_t*)c.mark, fc[1]);
  launch(c, "bc", dim3(nblocks(fc[2], 256)), dim3(256), 0, k_bc<2>, c.g, u[2], (const int8_t*)c.mark, fc[2]);
}

void run_extrapolate_bc(Ctx& c, bool recompute_phi) {
  if (recompute_phi) sweep_phi(c, c.phiair, AIR);
  sweep_u(c, c.au, c.phiair, AIR);
  if (recompute_phi) sweep_phi(c, c.phifluid, FLUID);
  sweep_u(c, c.wu, c.phifluid, FLUID);
  bc(c, c.au);
  bc(c, c.wu);
}

void run_divergence(Ctx& c) {
  const int64_t G = c.g.G;
  if (c.cfg.precision == MPF_FP64)
    launch(c, "divergence", dim3(nblocks(G, 256)), dim3(256), 0, k_divergence<double>, c.g, (const float*)c.wu[0],
           (const float*)c.wu[1], (const float*)c.wu[2], (const float*)c.au[0], (const float*)c.au[1],
           (const float*)c.au[2], (const int8_t*)c.mark, (const float*)c.lsmerge, c.cfg.solid_constraint,
           (double*)c.div);
  else
    launch(c, "divergence", dim3(nblocks(G, 256)), dim3(256), 0, k_divergence<float>, c.g, (const float*)c.wu[0],
           (const float*)c.wu[1], (const float*)c.wu[2], (const float*)c.au[0], (const float*)c.au[1],
           (const float*)c.au[2], (const int8_t*)c.mark, (const float*)c.lsmerge, c.cfg.solid_constraint,
           (float*)c.div);
}

template <class R>
static void subgrad_t(Ctx& c, float* const* u, double inv_rho) {
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  const int vd = c.cfg.op == MPF_OPERATOR_VARDENS, sc = c.cfg.solid_constraint;
  const R* p = (const R*)c.press;
  launch(c, "subgrad", dim3(nblocks(fc[0], 256)), dim3(256), 0, k_subgrad<0, R>, c.g, u[0], p, (const int8_t*)c.mark, vd,
         inv_rho, sc, fc[0]);
  launch(c, "subgrad", dim3(nblocks(fc[1], 256)), dim3(256), 0, k_subgrad<1, R>, c.g, u[1], p, (const int8_t*)c.mark, vd,
         inv_rho, sc, fc[1]);
  launch(c, "subgrad", dim3(nblocks(fc[2], 256)), dim3(256), 0, k_subgrad<2, R>, c.g, u[2], p, (const int8_t*)c.mark, vd,
         inv_rho, sc, fc[2]);
}

void run_subgrad(Ctx& c) {
  double il = 1.0 / (double)c.cfg.rho_liquid, ia = 1.0 / (double)c.cfg.rho_air;
  if (c.cfg.precision == MPF_FP32) {
    subgrad_t<float>(c, c.wu, il);
    subgrad_t<float>(c, c.au, ia);
  } else {
    subgrad_t<double>(c, c.wu, il);
    subgrad_t<double>(c, c.au, ia);
  }
}

void run_heat_grid_pre(Ctx& c) {
  const mpf_config& k = c.cfg;
  const int64_t G = c.g.G;
  const int ntx = (c.g.nx + ST_X - 1) / ST_X, nty = (c.g.ny + ST_Y - 1) / ST_Y, ntz = (c.g.nz + ST_Z - 1) / ST_Z;
  static_assert(HR <= ST_RTOT, "heat regions fit the P2G tile buffer");
  launch(c, "heat_p2g", dim3((unsigned)(ntx * nty * ntz)), dim3(ST_CELLS), 0, k_heat_tile, c.g,
         (const f3*)c.pos[c.cur], (const float*)c.temp[c.cur], (const uint32_t*)c.cstart, (const uint32_t*)c.cend,
         (const uint8_t*)c.cmask, c.p2g_parts, ntx, nty, ntz);
  launch(c, "heat_collect", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_collect, c.g, (const float*)c.p2g_parts, ntx,
         nty, ntz, k.default_temperature, c.tp);
  launch(c, "heat_pre", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_pre, c.g, c.tp, c.tsave, c.ctmp,
         (const int8_t*)c.mark, k.heat_add_bottom, k.heat_increase_bottom, k.heat_bottom_clamp, k.heat_alpha[0],
         k.heat_alpha[1], k.heat_alpha[2], k.heat_alpha[3], k.heat_alpha[4], k.heat_alpha[5], k.dt);
}

void run_heat_jacobi(Ctx& c, int iters) {
  const mpf_config& k = c.cfg;
  const int64_t G = c.g.G;
  float* x = (float*)c.vx;
  float* xn = (float*)c.vr;
  uint16_t* code = (uint16_t*)c.ctmp_u32;
  launch(c, "heat_start", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_start, G, (const float*)c.tsave,
         (const int8_t*)c.mark, x);
  launch(c, "heat_code", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_code, c.g, (const int8_t*)c.mark, code,
         k.heat_vacuum_neumann);
  for (int it = 0; it < iters; ++it) {
    launch(c, "heat_jacobi", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_jacobi, c.g, (const float*)x, xn,
           (const float*)c.ctmp, (const uint16_t*)code, k.heat_alpha[0], k.heat_alpha[1], k.heat_alpha[2],
           k.heat_alpha[3], k.heat_alpha[4], k.heat_alpha[5], k.dt);
    float* t = x; x = xn; xn = t;
  }
  hipMemcpyAsync(c.tp, x, (size_t)G * 4, hipMemcpyDeviceToDevice, c.stream);
}

void run_heat_grid_post(Ctx& c) {
  const int64_t G = c.g.G;
  hipMemcpyAsync(c.ctmp, c.tp, (size_t)G * 4, hipMemcpyDeviceToDevice, c.stream);
  launch(c, "heat_post", dim3(nblocks(G, 256)), dim3(256), 0, k_heat_post, c.g, (const float*)c.ctmp, c.tp,
         (const float*)c.tsave, c.tpchange, (const int8_t*)c.mark, c.cfg.heat_extend_boundary);
}

int run_marching_cubes(Ctx& c, mpf_mc_info* info) {
  static bool tables_loaded[64] = {};  // per device; the tables are constants
  const int dslot = c.device >= 0 && c.device < 64 ? c.device : 0;
  if (!tables_loaded[dslot]) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tri), MC_TRI_TABLE, sizeof(MC_TRI_TABLE)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_nv), MC_NUM_VERTS, sizeof(MC_NUM_VERTS)) != hipSuccess) {
      c.err = "marching cubes tables: hipMemcpyToSymbol failed";
      return MPF_HIP_ERROR;
    }
    tables_loaded[dslot] = true;
  }
  const int src = c.cfg.mc_source;
  McGrid m{};
  const float* f;
  if (src == MPF_MC_LEVELSET) {
    m.nx = c.g.nx; m.ny = c.g.ny; m.nz = c.g.nz;
    m.h = c.g.h;
    m.off = 0.5f;
    f = c.lsfluid;
  } else {
    m.nx = c.g2.nx + 1; m.ny = c.g2.ny + 1; m.nz = c.g2.nz + 1;
    m.h = c.cfg.dx / 2.0f;  // cellsize.x / (NXMC / NX), spray_k.cu:1640
    m.off = 0.0f;
    f = c.mc_field;
  }
  m.x0 = c.g.gx0; m.y0 = c.g.gy0; m.z0 = c.g.gz0;
  m.N = (int64_t)m.nx * m.ny * m.nz;
  m.V = (int64_t)(m.nx - 1) * (m.ny - 1) * (m.nz - 1);
  if (src != MPF_MC_LEVELSET) {
    const f3* pos = (const f3*)c.pos[c.cur];
    const int8_t* fl = (const int8_t*)c.flag[c.cur];
    const uint32_t *cs = c.mc_cstart, *ce = c.mc_cend;
    const dim3 gd(nblocks(m.N, 256));
    if (src == MPF_MC_LIQUID_GAS)
      launch(c, "mc_density_liquid_gas", gd, dim3(256), 0, k_mc_density<MPF_MC_LIQUID_GAS>, m, c.g2, pos, fl, cs, ce,
             c.mc_field, c.cfg.mc_gas_airsolo);
    else if (src == MPF_MC_GAS)
      launch(c, "mc_density_gas", gd, dim3(256), 0, k_mc_density<MPF_MC_GAS>, m, c.g2, pos, fl, cs, ce, c.mc_field,
             c.cfg.mc_gas_airsolo);
    else
      launch(c, "mc_density_solid", gd, dim3(256), 0, k_mc_density<MPF_MC_SOLID>, m, c.g2, pos, fl, cs, ce,
             c.mc_field, c.cfg.mc_gas_airsolo);
    c.mc_field_valid = true;
  }
  const int64_t V = m.V, N = m.N;
  c.mc_nvox = V;
  hipMemsetAsync(c.mc_edge, 0, (size_t)3 * N * 4, c.stream);
  uint32_t* spread = c.counters + 64;
  hipMemsetAsync(spread, 0, NSPREAD * SPREAD_STRIDE * 4, c.stream);
  launch(c, "mc_classify", dim3(nblocks(V, 256)), dim3(256), 0, k_mc_classify, m, f, c.mc_case, c.mc_tcount,
         c.mc_edge, spread);
  // vertex ids = exclusive scan of the used-edge flags (second half of mc_edge)
  uint32_t* vid = c.mc_edge + 3 * N;
  int64_t nv = run_scan_u32(c, c.mc_edge, vid, 3 * N, true);
  // triangle offsets = exclusive scan of per-voxel triangle counts (second half of mc_tcount)
  int64_t nt = run_scan_u32(c, c.mc_tcount, c.mc_tcount + V, V, true);
  if (nv < 0 || nt < 0) return MPF_HIP_ERROR;
  hipMemcpyAsync(c.counters_host + 64, spread, NSPREAD * SPREAD_STRIDE * 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "marching_cubes")) return MPF_HIP_ERROR;
  if (nv > c.cfg.mc_max_verts || nt > c.cfg.mc_max_tris) {
    c.err = "marching cubes mesh exceeds mc_max_verts/mc_max_tris";
    return MPF_CAPACITY;
  }
  launch(c, "mc_verts", dim3(nblocks(3 * N, 256)), dim3(256), 0, k_mc_verts, m, f, (const uint32_t*)c.mc_edge,
         (const uint32_t*)vid, c.mc_verts, c.mc_vedge);
  launch(c, "mc_tris", dim3(nblocks(V, 256)), dim3(256), 0, k_mc_tris, m, (const uint8_t*)c.mc_case,
         (const uint32_t*)(c.mc_tcount + V), (const uint32_t*)vid, c.mc_tris);
  // runMC_smooth step 8 (spray.cu:948-951): smoothing, then vertex normals
  // (the normal buffer is the smoothing's second vertex buffer before that)
  {
    const uint32_t* toff = c.mc_tcount + V;
    float* a = c.mc_verts;
    float* b = c.mc_normals;
    for (int it = 0; it < c.cfg.mc_smooth_iters; ++it) {
      launch(c, "mc_smooth", dim3(nblocks(nv, 256)), dim3(256), 0, k_mc_smooth, m, (const uint32_t*)c.mc_vedge,
             (const uint8_t*)c.mc_case, toff, (const uint32_t*)c.mc_tris, (const float*)a, b, nv, 0.5f);
      std::swap(a, b);
    }
    if (a != c.mc_verts && nv > 0)
      hipMemcpyAsync(c.mc_verts, a, (size_t)nv * 12, hipMemcpyDeviceToDevice, c.stream);
    launch(c, "mc_normals", dim3(nblocks(nv, 256)), dim3(256), 0, k_mc_normals, m, (const uint32_t*)c.mc_vedge,
           (const uint8_t*)c.mc_case, toff, (const uint32_t*)c.mc_tris, (const float*)c.mc_verts, c.mc_normals, nv);
  }
  c.mc_nv = nv;
  c.mc_nt = nt;
  if (info) {
    info->verts = nv;
    info->tris = nt;
    int64_t act = 0;
    for (int q = 0; q < NSPREAD; ++q) act += c.counters_host[64 + q * SPREAD_STRIDE];
    info->active_voxels = act;
  }
  return MPF_OK;
}

}  // namespace mpf
