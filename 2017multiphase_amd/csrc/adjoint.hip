// Hand-written adjoints (reverse-mode products) of the projection, P2G, G2P
// and advection stages.  The reference has no counterpart; these serve
// gradient computations through the step.  Each adjoint is evaluated at the
// state of the stage's forward call (flags, level sets, cell lists, particle
// order, positions) and holds every discrete decision of that call fixed.
// Cotangents are FP64.
#include "interp.h"
#include "mpf_internal.h"

namespace mpf {

namespace {

__device__ __forceinline__ int8_t mark_at(const Grid& g, const int8_t* mark, int i, int j, int k) {
  return valid(g, i, j, k) ? mark[cid(g, i, j, k)] : BOUNDARY;
}

struct Faces6 {
  double* f[6];  // water ux, uy, uz, air ux, uy, uz
};
struct CFaces6 {
  const double* f[6];
};

// ---------------- projection ----------------
// Forward (cspray::project_CG_bubble, spray.cu:1590-1611):
//   b = D(u_w, u_a) on FLUID|AIR cells (cptdivergence_bubble2, theta blend at
//   liquid/air faces; 0 at faces of constrained SOLID cells), p = A^-1 b,
//   u'_phase[f] = u_phase[f] - beta_phase (p[c1] - p[c0]) / h on interior faces
//   (subGradPress), 0 on faces touching a constrained SOLID cell.
// Adjoint: gp = -G^T g' (cell gather), lambda = A^-1 gp (A symmetric),
//   g_in = (identity part of g') + D^T lambda (face gather).

// gp[c] = sum over the 6 faces of c of (+-beta g'[f] / h), unknown cells only.
__global__ void k_adj_gradT(Grid g, const int8_t* __restrict__ mark, CFaces6 go, double bw, double ba, int solid_c,
                            double* __restrict__ gp) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const int8_t m = mark[idx];
  double s = 0.0;
  if (m == FLUID || m == AIR) {
    const int n3[3] = {g.nx, g.ny, g.nz};
    const int c3[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      // face below (c is its upper cell c1): u' -= beta (p[c] - p[below]) / h -> gp[c] -= beta g / h
      if (c3[a] > 0) {
        int o[3] = {0, 0, 0};
        o[a] = -1;
        bool solid = solid_c && (m == SOLID || mark_at(g, mark, i + o[0], j + o[1], k + o[2]) == SOLID);
        if (!solid) {
          int64_t f = fid(g, a, i, j, k);
          s -= (bw * go.f[a][f] + ba * go.f[3 + a][f]) / (double)g.h;
        }
      }
      // face above (c is its lower cell c0): gp[c] += beta g / h
      if (c3[a] + 1 < n3[a]) {
        int o[3] = {0, 0, 0};
        o[a] = 1;
        bool solid = solid_c && (m == SOLID || mark_at(g, mark, i + o[0], j + o[1], k + o[2]) == SOLID);
        if (!solid) {
          int64_t f = fid(g, a, i + o[0], j + o[1], k + o[2]);
          s += (bw * go.f[a][f] + ba * go.f[3 + a][f]) / (double)g.h;
        }
      }
    }
  }
  gp[idx] = s;
}

// Coefficients of face velocity (water, air) in the divergence of cell c for
// its face towards neighbour n (cptdivergence_bubble2 cases).
__device__ __forceinline__ void div_coef(const Grid& g, const int8_t* mark, const float* ls, int64_t c, int8_t m,
                                         int64_t n, int8_t mn, int solid_c, double& cw, double& ca) {
  cw = ca = 0.0;
  if (solid_c && mn == SOLID) return;
  if (m == FLUID && mn != AIR) { cw = 1.0; return; }
  if (m == AIR && mn != FLUID) { ca = 1.0; return; }
  double th = (0.0 - (double)ls[c]) / ((double)ls[n] - (double)ls[c]);
  if (m == FLUID) { cw = th; ca = 1.0 - th; }
  else { ca = th; cw = 1.0 - th; }
}

template <int A>
__global__ void k_adj_divT(Grid g, const int8_t* __restrict__ mark, const float* __restrict__ ls,
                           const double* __restrict__ lam, CFaces6 go, Faces6 gi, int solid_c, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const int n3[3] = {i, j, k};
  const int lim = (A == 0) ? g.nx : (A == 1) ? g.ny : g.nz;
  // cells on either side: c1 = (i,j,k) (upper), c0 = one step down along A
  const bool has1 = n3[A] < lim, has0 = n3[A] > 0;
  const int bi = i - (A == 0), bj = j - (A == 1), bk = k - (A == 2);
  const int8_t m1 = has1 ? mark[cid(g, i, j, k)] : BOUNDARY;
  const int8_t m0 = has0 ? mark[cid(g, bi, bj, bk)] : BOUNDARY;
  const bool interior = has0 && has1;
  const bool solid = solid_c && (m0 == SOLID || m1 == SOLID);
  double w = 0.0, a = 0.0;
  if (!(interior && solid)) {  // identity part (interior faces keep u; domain faces are unchanged)
    w = go.f[A][f];
    a = go.f[3 + A][f];
  }
  const double ih = 1.0 / (double)g.h;
  if (has1 && (m1 == FLUID || m1 == AIR)) {  // f is the lower face of c1: div -= u0 / h
    double cw, ca;
    const int64_t c1 = cid(g, i, j, k);
    div_coef(g, mark, ls, c1, m1, has0 ? cid(g, bi, bj, bk) : c1, m0, solid_c, cw, ca);
    w -= lam[c1] * cw * ih;
    a -= lam[c1] * ca * ih;
  }
  if (has0 && (m0 == FLUID || m0 == AIR)) {  // f is the upper face of c0: div += u1 / h
    double cw, ca;
    const int64_t c0 = cid(g, bi, bj, bk);
    div_coef(g, mark, ls, c0, m0, has1 ? cid(g, i, j, k) : c0, m1, solid_c, cw, ca);
    w += lam[c0] * cw * ih;
    a += lam[c0] * ca * ih;
  }
  gi.f[A][f] = w;
  gi.f[3 + A][f] = a;
}

// ---------------- P2G ----------------
// Forward (mapvelp2g_k_fluidSolid / _air, spray_k.cu:2809-2937):
//   u_f = sum_p w_pf v_p / W_f,  W_f = sum_p w_pf,  w_pf = m_p max(RE^2/max(r^2,1e-4) - 1, 0)
// Adjoint per particle p (gather over the faces whose window holds p's cell):
//   g_v[p] += (w_pf / W_f) g_u[f]
//   g_x[p] += (dw_pf/dx)(v_p - u_f) / W_f g_u[f]

// W_f and u_f recomputed in FP64 (face gather with the forward's windows).
template <int A, bool AIRPH>
__global__ void k_adj_p2g_faces(Grid g, const f3* __restrict__ pos, const f3* __restrict__ vel,
                                const float* __restrict__ mass, const int8_t* __restrict__ flag,
                                const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                                double* __restrict__ Wout, double* __restrict__ Uout, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const double gx = (double)i + (A == 0 ? 0.0 : 0.5), gy = (double)j + (A == 1 ? 0.0 : 0.5),
               gz = (double)k + (A == 2 ? 0.0 : 0.5);
  constexpr int lo = AIRPH ? -2 : -1, hin = AIRPH ? 1 : 0, lot = AIRPH ? -2 : -1, hit = AIRPH ? 2 : 1;
  const double RE2 = (double)1.4f * (double)1.4f, scale = 1.0 / (double)g.h;
  double W = 0, U = 0;
  for (int di = (A == 0 ? lo : lot); di <= (A == 0 ? hin : hit); ++di)
    for (int dj = (A == 1 ? lo : lot); dj <= (A == 1 ? hin : hit); ++dj)
      for (int dk = (A == 2 ? lo : lot); dk <= (A == 2 ? hin : hit); ++dk) {
        if (!valid(g, i + di, j + dj, k + dk)) continue;
        int64_t c = cid(g, i + di, j + dj, k + dk);
        uint32_t s = cstart[c];
        if (s == MPF_CELL_UNDEF) continue;
        for (uint32_t p = s; p < cend[c]; ++p) {
          int8_t fl = flag[p];
          bool take = AIRPH ? (fl == AIR) : (fl == FLUID || fl == SOLID);
          if (!take) continue;
          f3 x = pos[p];
          double dx = ((double)x.x - g.gx0) * scale - gx, dy = ((double)x.y - g.gy0) * scale - gy,
                 dz = ((double)x.z - g.gz0) * scale - gz;
          double r2 = dx * dx + dy * dy + dz * dz;
          double w = (double)mass[p] * fmax(RE2 / fmax(r2, (double)0.0001f) - 1.0, 0.0);
          W += w;
          const float* vp = &vel[p].x;
          U += w * (double)vp[A];
        }
      }
  Wout[f] = W;
  Uout[f] = (W > 0) ? U / W : 0.0;
}

struct P2GFaceData {
  const double* W[3];
  const double* U[3];
  const double* gu[3];
};

template <bool AIRPH>
__global__ void k_adj_p2g_part(Grid g, int64_t P, const f3* __restrict__ pos, const f3* __restrict__ vel,
                               const float* __restrict__ mass, const int8_t* __restrict__ flag, P2GFaceData fd,
                               double* __restrict__ gv, double* __restrict__ gx) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  const bool take = AIRPH ? (fl == AIR) : (fl == FLUID || fl == SOLID);
  if (!take) return;
  const f3 x = pos[p];
  const float* vp = &vel[p].x;
  // the particle's cell, as the hash assigns it (getijkfrompos)
  float qx = (x.x - g.gx0) / g.h, qy = (x.y - g.gy0) / g.h, qz = (x.z - g.gz0) / g.h;
  const int ci = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
  const int cj = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
  const int ck = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
  const double RE2 = (double)1.4f * (double)1.4f, scale = 1.0 / (double)g.h;
  const double X = ((double)x.x - g.gx0) * scale, Y = ((double)x.y - g.gy0) * scale, Z = ((double)x.z - g.gz0) * scale;
  const double m = (double)mass[p];
  constexpr int lo = AIRPH ? -2 : -1, hin = AIRPH ? 1 : 0, lot = AIRPH ? -2 : -1, hit = AIRPH ? 2 : 1;
  double av[3] = {0, 0, 0}, ax[3] = {0, 0, 0};
#pragma unroll
  for (int A = 0; A < 3; ++A) {
    const int dims[3] = {g.nx + (A == 0), g.ny + (A == 1), g.nz + (A == 2)};
    // faces f = c - o for window offsets o
    for (int oi = (A == 0 ? lo : lot); oi <= (A == 0 ? hin : hit); ++oi)
      for (int oj = (A == 1 ? lo : lot); oj <= (A == 1 ? hin : hit); ++oj)
        for (int ok = (A == 2 ? lo : lot); ok <= (A == 2 ? hin : hit); ++ok) {
          const int fi = ci - oi, fj = cj - oj, fk = ck - ok;
          if (fi < 0 || fj < 0 || fk < 0 || fi >= dims[0] || fj >= dims[1] || fk >= dims[2]) continue;
          const int64_t f = fid(g, A, fi, fj, fk);
          const double W = fd.W[A][f];
          if (!(W > 0)) continue;
          const double gxf = (double)fi + (A == 0 ? 0.0 : 0.5), gyf = (double)fj + (A == 1 ? 0.0 : 0.5),
                       gzf = (double)fk + (A == 2 ? 0.0 : 0.5);
          const double dx = X - gxf, dy = Y - gyf, dz = Z - gzf;
          const double r2 = dx * dx + dy * dy + dz * dz;
          const double r2c = fmax(r2, (double)0.0001f);
          const double s = RE2 / r2c - 1.0;
          if (!(s > 0)) continue;
          const double w = m * s;
          const double gu = fd.gu[A][f];
          av[A] += (w / W) * gu;
          if (r2 > (double)0.0001f) {
            // dw/dx_d = -2 m RE^2 d_d / r^4 * scale
            const double dwr = -2.0 * m * RE2 / (r2 * r2) * scale;
            const double coef = ((double)vp[A] - fd.U[A][f]) / W * gu * dwr;
            ax[0] += coef * dx;
            ax[1] += coef * dy;
            ax[2] += coef * dz;
          }
        }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    gv[3 * p + d] += av[d];
    gx[3 * p + d] += ax[d];
  }
}

// ---------------- G2P ----------------
// Forward (mapvelg2p_flip_bubble spray_k.cu:3525-3540 with blend alpha):
//   v' = (1 - alpha) T(x) u + alpha (v + T(x) du),  du = u - u_old (grid of the particle's phase)
// Adjoint: g_v = alpha g_v';  g_u += T^T g_v' (both terms);  g_uold -= alpha T^T g_v';
//   g_x += (1 - alpha) (dT u/dx)^T g_v' + alpha (dT du/dx)^T g_v'.
// The grid scatter uses FP64 atomics (order-dependent rounding).
struct G2PArgs {
  const float* u[6];   // water u, air u (current)
  const float* du[6];  // water du, air du (forward's u - u_old)
  double* gu[6];
  double* guo[6];
};

__global__ void k_adj_g2p(Grid g, int64_t P, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                          G2PArgs a, double alpha, const double* __restrict__ gvo, double* __restrict__ gvi,
                          double* __restrict__ gx) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  int ph = -1;
  if (fl == FLUID || fl == SOLID) ph = 0;
  else if (fl == AIR) ph = 1;
  const double gv[3] = {gvo[3 * p], gvo[3 * p + 1], gvo[3 * p + 2]};
  if (ph < 0) {
    for (int d = 0; d < 3; ++d) gvi[3 * p + d] += gv[d];
    return;
  }
  for (int d = 0; d < 3; ++d) gvi[3 * p + d] += alpha * gv[d];
  const f3 x = pos[p];
  const double X = ((double)x.x - g.gx0) / g.h, Y = ((double)x.y - g.gy0) / g.h, Z = ((double)x.z - g.gz0) / g.h;
  double gpos[3] = {0, 0, 0};
#pragma unroll
  for (int A = 0; A < 3; ++A) {
    const int w = g.nx + (A == 0), hh = g.ny + (A == 1), d = g.nz + (A == 2);
    const double qx = X - (A == 0 ? 0.0 : 0.5), qy = Y - (A == 1 ? 0.0 : 0.5), qz = Z - (A == 2 ? 0.0 : 0.5);
    int i, j, k;
    double wt[8];
    trilinear_weights(qx, qy, qz, w, hh, d, i, j, k, wt);
    const int64_t sj = d, si = (int64_t)hh * d;
    const int64_t base = i * si + j * sj + k;
    const int64_t off[8] = {0, si, sj, si + sj, 1, si + 1, sj + 1, si + sj + 1};
    double* gu = a.gu[3 * ph + A];
    double* guo = a.guo[3 * ph + A];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (wt[c] == 0.0) continue;
      atomicAdd(&gu[base + off[c]], wt[c] * gv[A]);
      if (alpha != 0.0) atomicAdd(&guo[base + off[c]], -alpha * wt[c] * gv[A]);
    }
    double g1x, g1y, g1z, g2x, g2y, g2z;
    trilinear_grad(a.u[3 * ph + A], qx, qy, qz, w, hh, d, g1x, g1y, g1z);
    trilinear_grad(a.du[3 * ph + A], qx, qy, qz, w, hh, d, g2x, g2y, g2z);
    const double s1 = (1.0 - alpha) * gv[A], s2 = alpha * gv[A];
    gpos[0] += (s1 * g1x + s2 * g2x) / (double)g.h;
    gpos[1] += (s1 * g1y + s2 * g2y) / (double)g.h;
    gpos[2] += (s1 * g1z + s2 * g2z) / (double)g.h;
  }
  for (int d = 0; d < 3; ++d) gx[3 * p + d] += gpos[d];
}

// ---------------- advection ----------------
// Forward (CollisionWithSolid_k, spray_k.cu:5276-5345, away from solids):
//   x' = clamp(x + v dt), v' = v with clamped components zeroed; SOLID fixed.
// Adjoint, per component not clamped: g_x = g_x', g_v = g_v' + dt g_x';
// clamped component: both 0.  The solid-collision response (phi_solid <= 0.5)
// is not differentiated; particles in that band are reported through
// `in_band` (count) so a caller can tell.
__global__ void k_adj_advect(Grid g, int64_t P, const f3* __restrict__ pos, const f3* __restrict__ vel,
                             const int8_t* __restrict__ flag, const float* __restrict__ phis, double dt,
                             const double* __restrict__ gxo, const double* gvo, double* __restrict__ gxi,
                             double* gvi, uint32_t* __restrict__ in_band) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  if (fl == SOLID) {
    for (int d = 0; d < 3; ++d) {
      gxi[3 * p + d] = gxo[3 * p + d];
      gvi[3 * p + d] = gvo[3 * p + d];
    }
    return;
  }
  const f3 x = pos[p], v = vel[p];
  if (phis && scalar_from_grid(g, x, phis) <= 0.5f) atomicAdd(in_band, 1u);
  const float rate = (fl == AIRSOLO) ? 0.8f : 0.5f;
  const float m = g.h + rate * g.h;
  const float lo[3] = {g.gx0 + m, g.gy0 + m, g.gz0 + m}, hi[3] = {g.gx1 - m, g.gy1 - m, g.gz1 - m};
  const float xs[3] = {x.x + v.x * (float)dt, x.y + v.y * (float)dt, x.z + v.z * (float)dt};
  for (int d = 0; d < 3; ++d) {
    const bool cl = xs[d] <= lo[d] || xs[d] >= hi[d];
    gxi[3 * p + d] = cl ? 0.0 : gxo[3 * p + d];
    gvi[3 * p + d] = cl ? 0.0 : gvo[3 * p + d] + dt * gxo[3 * p + d];
  }
}

// ---------------- extrapolation and wall rule ----------------
// Forward (sweepU + setWaterBoundaryU_k, spray_k.cu:900-995): two passes of
// out[f] = u[f], or, for a face with both cells outside phase sf, the blend
// wx u[f+dx e_x] + wy u[f+dy e_y] + wz u[f+dz e_z] of the last qualifying
// diagonal direction; then the wall rule zeroes faces and sets the ceiling
// uz(nz-1) to min(uz(nz-2), 0).  Linear in u for fixed level set and flags.

// The forward's decision for face (i,j,k) of axis A: false = identity.
template <int A>
__device__ __forceinline__ bool sweep_rule(const Grid& g, const float* __restrict__ ls,
                                           const int8_t* __restrict__ mark, int8_t sf, int i, int j, int k,
                                           int (&dsel)[3], float (&wsel)[3]) {
  const int n[3] = {i, j, k};
  const int lim = (A == 0) ? g.nx : (A == 1) ? g.ny : g.nz;
  if (!(n[A] > 1 && n[A] < lim - 1)) return false;
  const int bi = i - (A == 0), bj = j - (A == 1), bk = k - (A == 2);
  if (mark[cid(g, i, j, k)] == sf || mark[cid(g, bi, bj, bk)] == sf) return false;
  const float lsc = ls[cid(g, i, j, k)], lsb = ls[cid(g, bi, bj, bk)];
  constexpr int t1 = (A == 0) ? 1 : (A == 1) ? 0 : 1;
  constexpr int t2 = (A == 0) ? 2 : (A == 1) ? 2 : 0;
  const int lims[3] = {g.nx, g.ny, g.nz};
  bool any = false;
  for (int di = -1; di <= 1; di += 2)
    for (int dj = -1; dj <= 1; dj += 2)
      for (int dk = -1; dk <= 1; dk += 2) {
        const int dd[3] = {di, dj, dk};
        if (n[t1] + dd[t1] < 0 || n[t1] + dd[t1] > lims[t1] - 1) continue;
        if (n[t2] + dd[t2] < 0 || n[t2] + dd[t2] > lims[t2] - 1) continue;
        float wv[3];
        wv[A] = (float)(-dd[A]) * (lsc - lsb);
        if (wv[A] < 0) continue;
        int o1[3] = {0, 0, 0};
        o1[t1] = dd[t1];
        wv[t1] = (lsc + lsb - ls[cid(g, i + o1[0], j + o1[1], k + o1[2])] - ls[cid(g, bi + o1[0], bj + o1[1], bk + o1[2])]) * 0.5f;
        if (wv[t1] < 0) continue;
        int o2[3] = {0, 0, 0};
        o2[t2] = dd[t2];
        wv[t2] = (lsc + lsb - ls[cid(g, i + o2[0], j + o2[1], k + o2[2])] - ls[cid(g, bi + o2[0], bj + o2[1], bk + o2[2])]) * 0.5f;
        if (wv[t2] < 0) continue;
        float wx = wv[0], wy = wv[1], wz = wv[2];
        const float wsum = wx + wy + wz;
        if (wsum == 0) wx = wy = wz = 1.0f / 3;
        else { wx /= wsum; wy /= wsum; wz /= wsum; }
        dsel[0] = di; dsel[1] = dj; dsel[2] = dk;
        wsel[0] = wx; wsel[1] = wy; wsel[2] = wz;
        any = true;
      }
  return any;
}

// One transposed pass, as a gather: face n collects its identity term and the
// weights of the (at most 6) faces whose chosen blend reads it.
template <int A>
__global__ void k_adj_sweep_u(Grid g, const double* __restrict__ gout, double* __restrict__ gin,
                              const float* __restrict__ ls, const int8_t* __restrict__ mark, int8_t sf, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const int dims[3] = {g.nx + (A == 0), g.ny + (A == 1), g.nz + (A == 2)};
  int dsel[3];
  float wsel[3];
  double acc = sweep_rule<A>(g, ls, mark, sf, i, j, k, dsel, wsel) ? 0.0 : gout[f];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int s = -1; s <= 1; s += 2) {
      int q[3] = {i, j, k};
      q[d] -= s;  // candidate reader: its blend along d with direction s lands on f
      if (q[d] < 0 || q[d] >= dims[d]) continue;
      if (!sweep_rule<A>(g, ls, mark, sf, q[0], q[1], q[2], dsel, wsel)) continue;
      if (dsel[d] != s) continue;
      acc += (double)wsel[d] * gout[fid(g, A, q[0], q[1], q[2])];
    }
  gin[f] = acc;
}

__device__ __forceinline__ bool isb(const Grid& g, const int8_t* mark, int i, int j, int k) {
  return mark[cid(g, i, j, k)] == BOUNDARY;
}
// 0 = kept, 1 = zeroed, 2 = ceiling face (k = nz-1 of uz); same order of tests as the forward.
template <int A>
__device__ __forceinline__ int bc_rule(const Grid& g, const int8_t* mark, int i, int j, int k) {
  if (A == 0) {
    if (i <= 1 || i >= (g.nx + 1) - 2) return 1;
    if (i > 1 && i < g.nx - 1 && (isb(g, mark, i, j, k) != isb(g, mark, i - 1, j, k))) return 1;
    return 0;
  } else if (A == 1) {
    if (j <= 1 || j >= (g.ny + 1) - 2) return 1;
    if (j > 0 && j < g.ny && (isb(g, mark, i, j, k) != isb(g, mark, i, j - 1, k))) return 1;
    return 0;
  } else {
    const int zn = g.nz + 1;
    if (k <= 1 || k >= zn - 1) return 1;
    if (k == zn - 2) return 2;
    if (k > 0 && k < g.nz && (isb(g, mark, i, j, k) != isb(g, mark, i, j, k - 1))) return 1;
    return 0;
  }
}

// Transposed wall rule.  The ceiling's branch (uz(nz-2) < 0) is read from the
// forward's output `u_after`: the output is negative exactly when it copied.
template <int A>
__global__ void k_adj_bc(Grid g, const double* __restrict__ gout, double* __restrict__ gin,
                         const int8_t* __restrict__ mark, const float* __restrict__ u_after, int64_t nf) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const int r = bc_rule<A>(g, mark, i, j, k);
  double acc = (r == 0) ? gout[f] : 0.0;
  if (A == 2 && r == 0 && k + 1 <= g.nz && bc_rule<2>(g, mark, i, j, k + 1) == 2 && u_after[f + 1] < 0.0f)
    acc += gout[f + 1];
  gin[f] = acc;
}

__global__ void k_fill_d(int64_t n, double* __restrict__ a, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace

static int64_t face_count(const Ctx& c, int q) {
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  return fc[q % 3];
}

// Adjoint scratch, allocated by mpf_create when cfg.adjoint_enabled is set
// (a failure part-way leaves the pointers set so that adj_free releases them).
bool adj_alloc_all(Ctx& c) {
  auto get = [&](double*& p, size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return false;
    p = (double*)q;
    return true;
  };
  for (int q = 0; q < 6; ++q)
    if (!get(c.adj_fin[q], (size_t)face_count(c, q) * 8) || !get(c.adj_fout[q], (size_t)face_count(c, q) * 8))
      return false;
  for (int q = 0; q < 2; ++q)
    if (!get(c.adj_cell[q], (size_t)c.g.G * 8)) return false;
  for (int q = 0; q < 3; ++q)
    if (!get(c.adj_part[q], (size_t)(c.cap > 0 ? c.cap : 1) * 24)) return false;
  c.adj_ready = true;
  return true;
}

static bool adj_alloc(Ctx& c) {
  if (c.adj_ready) return true;
  c.err = "adjoint entry points need mpf_config.adjoint_enabled = 1 at mpf_create";
  return false;
}

void adj_free(Ctx& c) {
  for (int q = 0; q < 6; ++q) {
    if (c.adj_fin[q]) hipFree(c.adj_fin[q]);
    if (c.adj_fout[q]) hipFree(c.adj_fout[q]);
    c.adj_fin[q] = c.adj_fout[q] = nullptr;
  }
  for (int q = 0; q < 2; ++q) {
    if (c.adj_cell[q]) hipFree(c.adj_cell[q]);
    c.adj_cell[q] = nullptr;
  }
  for (int q = 0; q < 3; ++q) {
    if (c.adj_part[q]) hipFree(c.adj_part[q]);
    c.adj_part[q] = nullptr;
  }
  c.adj_ready = false;
}

static void copy_in(Ctx& c, double* dst, const double* src, int64_t n, int where) {
  if (src) hipMemcpyAsync(dst, src, (size_t)n * 8, where == MPF_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                          c.stream);
  else launch(c, "adj_fill", dim3(nblocks(n, 256)), dim3(256), 0, k_fill_d, n, dst, 0.0);
}
static void copy_out(Ctx& c, double* dst, const double* src, int64_t n, int where) {
  if (dst) hipMemcpyAsync(dst, src, (size_t)n * 8, where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                          c.stream);
}
static void zero(Ctx& c, double* a, int64_t n) {
  launch(c, "adj_fill", dim3(nblocks(n, 256)), dim3(256), 0, k_fill_d, n, a, 0.0);
}

int run_project_adjoint(Ctx& c, const double* const* gout, double* const* gin, int where, mpf_solve_info* info) {
  if (!adj_alloc(c)) return MPF_INVALID_ARG;
  for (int q = 0; q < 6; ++q) copy_in(c, c.adj_fin[q], gout ? gout[q] : nullptr, face_count(c, q), where);
  const bool vd = c.cfg.op == MPF_OPERATOR_VARDENS;
  const double bw = vd ? 1.0 / (double)c.cfg.rho_liquid : 1.0, ba = vd ? 1.0 / (double)c.cfg.rho_air : 1.0;
  CFaces6 go;
  Faces6 gi;
  for (int q = 0; q < 6; ++q) {
    go.f[q] = c.adj_fin[q];
    gi.f[q] = c.adj_fout[q];
  }
  const int64_t G = c.g.G;
  launch(c, "adj_gradT", dim3(nblocks(G, 256)), dim3(256), 0, k_adj_gradT, c.g, (const int8_t*)c.mark, go, bw, ba,
         c.cfg.solid_constraint, c.adj_cell[0]);
  int st = run_adjoint_solve(c, c.adj_cell[0], c.adj_cell[1], info);
  if (st == MPF_HIP_ERROR || st == MPF_NONFINITE) return st;
  launch(c, "adj_divT", dim3(nblocks(c.g.FX, 256)), dim3(256), 0, k_adj_divT<0>, c.g, (const int8_t*)c.mark,
         (const float*)c.lsmerge, (const double*)c.adj_cell[1], go, gi, c.cfg.solid_constraint, c.g.FX);
  launch(c, "adj_divT", dim3(nblocks(c.g.FY, 256)), dim3(256), 0, k_adj_divT<1>, c.g, (const int8_t*)c.mark,
         (const float*)c.lsmerge, (const double*)c.adj_cell[1], go, gi, c.cfg.solid_constraint, c.g.FY);
  launch(c, "adj_divT", dim3(nblocks(c.g.FZ, 256)), dim3(256), 0, k_adj_divT<2>, c.g, (const int8_t*)c.mark,
         (const float*)c.lsmerge, (const double*)c.adj_cell[1], go, gi, c.cfg.solid_constraint, c.g.FZ);
  for (int q = 0; q < 6; ++q) copy_out(c, gin[q], c.adj_fout[q], face_count(c, q), where);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "project_adjoint")) return MPF_HIP_ERROR;
  return st;
}

int run_extrapolate_adjoint(Ctx& c, const double* const* gout, double* const* gin, int where) {
  if (!adj_alloc(c)) return MPF_INVALID_ARG;
  for (int q = 0; q < 6; ++q) copy_in(c, c.adj_fin[q], gout ? gout[q] : nullptr, face_count(c, q), where);
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  for (int ph = 0; ph < 2; ++ph) {
    float* const* u = ph == 0 ? c.wu : c.au;
    const float* phi = ph == 0 ? c.phifluid : c.phiair;
    const int8_t sf = ph == 0 ? FLUID : AIR;
    double* a = nullptr;
    double* b = nullptr;
    // wall rule^T: fin -> fout; then two sweep passes^T: fout -> fin -> fout
    for (int A = 0; A < 3; ++A) {
      a = c.adj_fin[3 * ph + A];
      b = c.adj_fout[3 * ph + A];
      const unsigned nb = nblocks(fc[A], 256);
      const int8_t* mk = c.mark;
      if (A == 0) launch(c, "adj_bc", dim3(nb), dim3(256), 0, k_adj_bc<0>, c.g, (const double*)a, b, mk, (const float*)u[0], fc[0]);
      if (A == 1) launch(c, "adj_bc", dim3(nb), dim3(256), 0, k_adj_bc<1>, c.g, (const double*)a, b, mk, (const float*)u[1], fc[1]);
      if (A == 2) launch(c, "adj_bc", dim3(nb), dim3(256), 0, k_adj_bc<2>, c.g, (const double*)a, b, mk, (const float*)u[2], fc[2]);
      for (int pass = 0; pass < 2; ++pass) {
        const double* src = pass == 0 ? b : a;
        double* dst = pass == 0 ? a : b;
        if (A == 0) launch(c, "adj_sweep_u", dim3(nb), dim3(256), 0, k_adj_sweep_u<0>, c.g, src, dst, phi, mk, sf, fc[0]);
        if (A == 1) launch(c, "adj_sweep_u", dim3(nb), dim3(256), 0, k_adj_sweep_u<1>, c.g, src, dst, phi, mk, sf, fc[1]);
        if (A == 2) launch(c, "adj_sweep_u", dim3(nb), dim3(256), 0, k_adj_sweep_u<2>, c.g, src, dst, phi, mk, sf, fc[2]);
      }
    }
  }
  for (int q = 0; q < 6; ++q) copy_out(c, gin[q], c.adj_fout[q], face_count(c, q), where);
  return hip_ok(c, hipStreamSynchronize(c.stream), "extrapolate_adjoint") ? MPF_OK : MPF_HIP_ERROR;
}

int run_p2g_adjoint(Ctx& c, const double* const* gfaces, double* gvel, double* gpos, int where) {
  if (!adj_alloc(c)) return MPF_INVALID_ARG;
  const int cur = c.cur;
  const int64_t P = c.P;
  for (int q = 0; q < 6; ++q) copy_in(c, c.adj_fin[q], gfaces ? gfaces[q] : nullptr, face_count(c, q), where);
  zero(c, c.adj_part[0], 3 * P);
  zero(c, c.adj_part[2], 3 * P);
  for (int ph = 0; ph < 2; ++ph) {
    // W in adj_fout[0..2], U in adj_fout[3..5]
    const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
    for (int A = 0; A < 3; ++A) {
      auto args = [&](auto kern) {
        launch(c, "adj_p2g_faces", dim3(nblocks(fc[A], 256)), dim3(256), 0, kern, c.g, (const f3*)c.pos[cur],
               (const f3*)c.vel[cur], (const float*)c.mass[cur], (const int8_t*)c.flag[cur],
               (const uint32_t*)c.cstart, (const uint32_t*)c.cend, c.adj_fout[A], c.adj_fout[3 + A], fc[A]);
      };
      if (ph == 0) {
        if (A == 0) args(k_adj_p2g_faces<0, false>);
        if (A == 1) args(k_adj_p2g_faces<1, false>);
        if (A == 2) args(k_adj_p2g_faces<2, false>);
      } else {
        if (A == 0) args(k_adj_p2g_faces<0, true>);
        if (A == 1) args(k_adj_p2g_faces<1, true>);
        if (A == 2) args(k_adj_p2g_faces<2, true>);
      }
    }
    P2GFaceData fd;
    for (int A = 0; A < 3; ++A) {
      fd.W[A] = c.adj_fout[A];
      fd.U[A] = c.adj_fout[3 + A];
      fd.gu[A] = c.adj_fin[3 * ph + A];
    }
    if (ph == 0)
      launch(c, "adj_p2g_part", dim3(nblocks(P, 256)), dim3(256), 0, k_adj_p2g_part<false>, c.g, P,
             (const f3*)c.pos[cur], (const f3*)c.vel[cur], (const float*)c.mass[cur], (const int8_t*)c.flag[cur], fd,
             c.adj_part[0], c.adj_part[2]);
    else
      launch(c, "adj_p2g_part", dim3(nblocks(P, 256)), dim3(256), 0, k_adj_p2g_part<true>, c.g, P,
             (const f3*)c.pos[cur], (const f3*)c.vel[cur], (const float*)c.mass[cur], (const int8_t*)c.flag[cur], fd,
             c.adj_part[0], c.adj_part[2]);
  }
  copy_out(c, gvel, c.adj_part[0], 3 * P, where);
  copy_out(c, gpos, c.adj_part[2], 3 * P, where);
  return hip_ok(c, hipStreamSynchronize(c.stream), "p2g_adjoint") ? MPF_OK : MPF_HIP_ERROR;
}

int run_g2p_adjoint(Ctx& c, const double* gvel_out, double* gvel_in, double* const* gu, double* const* guo,
                    double* gpos, int where) {
  if (!adj_alloc(c)) return MPF_INVALID_ARG;
  const int cur = c.cur;
  const int64_t P = c.P;
  copy_in(c, c.adj_part[0], gvel_out, 3 * P, where);
  zero(c, c.adj_part[1], 3 * P);
  zero(c, c.adj_part[2], 3 * P);
  for (int q = 0; q < 6; ++q) {
    zero(c, c.adj_fin[q], face_count(c, q));
    zero(c, c.adj_fout[q], face_count(c, q));
  }
  G2PArgs a;
  for (int A = 0; A < 3; ++A) {
    a.u[A] = c.wu[A];
    a.u[3 + A] = c.au[A];
    a.du[A] = c.wuo[A];
    a.du[3 + A] = c.auo[A];
  }
  for (int q = 0; q < 6; ++q) {
    a.gu[q] = c.adj_fin[q];
    a.guo[q] = c.adj_fout[q];
  }
  launch(c, "adj_g2p", dim3(nblocks(P, 256)), dim3(256), 0, k_adj_g2p, c.g, P, (const f3*)c.pos[cur],
         (const int8_t*)c.flag[cur], a, (double)c.cfg.flip_alpha, (const double*)c.adj_part[0], c.adj_part[1],
         c.adj_part[2]);
  copy_out(c, gvel_in, c.adj_part[1], 3 * P, where);
  copy_out(c, gpos, c.adj_part[2], 3 * P, where);
  for (int q = 0; q < 6; ++q) {
    copy_out(c, gu ? gu[q] : nullptr, c.adj_fin[q], face_count(c, q), where);
    copy_out(c, guo ? guo[q] : nullptr, c.adj_fout[q], face_count(c, q), where);
  }
  return hip_ok(c, hipStreamSynchronize(c.stream), "g2p_adjoint") ? MPF_OK : MPF_HIP_ERROR;
}

int run_advect_adjoint(Ctx& c, const double* gpos_out, const double* gvel_out, double* gpos_in, double* gvel_in,
                       int where, int64_t* in_band) {
  if (!adj_alloc(c)) return MPF_INVALID_ARG;
  if (c.cfg.integrator != MPF_ADVECT_EULER_COLLIDE) {
    c.err = "advection adjoint is implemented for the Euler/collision integrator";
    return MPF_INVALID_ARG;
  }
  const int cur = c.cur;
  const int64_t P = c.P;
  copy_in(c, c.adj_part[0], gvel_out, 3 * P, where);
  copy_in(c, c.adj_part[2], gpos_out, 3 * P, where);
  // the solid level set of this step's forward call (recomputed from the cell
  // lists, which the entry point requires to be current)
  run_solid_phi_and_velocity(c);
  hipMemsetAsync(c.counters + 4, 0, 4, c.stream);
  // g_x -> adj_part[1]; g_v is updated in place in adj_part[0] (each thread
  // reads its own entries before writing them)
  launch(c, "adj_advect", dim3(nblocks(P, 256)), dim3(256), 0, k_adj_advect, c.g, P, (const f3*)c.pos[cur],
         (const f3*)c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.phisolid, (double)c.cfg.dt,
         (const double*)c.adj_part[2], (const double*)c.adj_part[0], c.adj_part[1], c.adj_part[0],
         c.counters + 4);
  copy_out(c, gpos_in, c.adj_part[1], 3 * P, where);
  copy_out(c, gvel_in, c.adj_part[0], 3 * P, where);
  hipMemcpyAsync(c.counters_host + 4, c.counters + 4, 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "advect_adjoint")) return MPF_HIP_ERROR;
  if (in_band) *in_band = c.counters_host[4];
  return MPF_OK;
}

}  // namespace mpf
