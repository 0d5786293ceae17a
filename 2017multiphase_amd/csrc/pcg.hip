// Matrix-free (P)CG for the pressure projection (a12) and the implicit heat
// diffusion (a16).
//
// Replaces solver_cg_bubble (spray.cu:1613-1687) with computeAx_bubble /
// pcg_op_bubble / arrayproduct_k (spray_k.cu:3423-3460, 2279-2297) and
// solver_cg_heat (spray.cu:1780-1852) with computeAx_heat / pcg_op_heat
// (spray_k.cu:4526-4582).
//
// The scalar recurrence (alpha, beta, convergence) stays on the device: each
// kernel reduces its dot products into per-block partials; the last block to
// arrive (arrival counter + fences) sums the partials in block order —
// deterministic — and updates the SolveScalars record the next kernel reads.
// The host only polls the record every check_every iterations.
//
// Per iteration: spmv (q = A p, p.q) -> update (x, r, z = r/diag, r.z, r.r)
// -> [MIXED: fold + FP64 true residual when due] -> direction (p = z + beta p).
// Blocks own contiguous chunks of cells, chunk order spread so each XCD (the
// hardware dispatches blocks round-robin over the 8 XCDs) walks one
// contiguous slab and keeps stencil neighbours in its own L2.
#include "mpf_internal.h"

namespace mpf {

namespace {

constexpr int PB = 256;        // threads per block
constexpr int MAX_BLK = 2048;  // blocks per solver launch
constexpr int OP_UNIT = 0, OP_VARDENS = 1, OP_HEAT = 2;

struct OpArgs {
  Grid g;
  const int8_t* mark;
  const float* ls;
  double il, ia;    // 1/rho_liquid, 1/rho_air
  double hal[6];    // h^2/(alpha*dt) per cell type (heat)
  int solid_c;      // SOLID cells Neumann (pressure)
  int vac_neumann;  // VACUUM neighbours Neumann (heat)
};

struct Launch {
  int nblk;
  int64_t chunk;
};

__device__ __forceinline__ bool is_unknown(int OPk, int8_t m) {
  return OPk == OP_HEAT ? (m != BOUNDARY) : (m == FLUID || m == AIR);
}
__device__ __forceinline__ int8_t mark_at(const Grid& g, const int8_t* mark, int i, int j, int k) {
  return valid(g, i, j, k) ? mark[cid(g, i, j, k)] : BOUNDARY;
}

// 1/rho of the face between cells ca (an unknown) and cb; FLUID|AIR faces
// use theta of the lower-index cell so the operator is exactly symmetric.
template <class A>
__device__ __forceinline__ A face_beta(const OpArgs& o, int64_t ca, int64_t cb, int8_t ma, int8_t mb) {
  const A il = (A)o.il, ia = (A)o.ia;
  bool ub = (mb == FLUID || mb == AIR);
  if (!ub || ma == mb) return ma == FLUID ? il : ia;
  int64_t lo = ca < cb ? ca : cb, hi = ca < cb ? cb : ca;
  int8_t mlo = ca < cb ? ma : mb, mhi = ca < cb ? mb : ma;
  A ll = (A)o.ls[lo], lh = (A)o.ls[hi];
  A th = -ll / (lh - ll);
  th = th < A(0) ? A(0) : (th > A(1) ? A(1) : th);
  A blo = mlo == FLUID ? il : ia, bhi = mhi == FLUID ? il : ia;
  return bhi + th * (blo - bhi);
}

// (A x)[idx] in arithmetic A, x stored as X.  Neighbour order +x,+y,+z,-x,-y,-z
// as computeAx_bubble.
template <class A, int OPk, class X>
__device__ __forceinline__ A apply_op(const OpArgs& o, const X* __restrict__ x, int64_t idx) {
  const Grid& g = o.g;
  const int8_t m = o.mark[idx];
  if (!is_unknown(OPk, m)) return A(0);
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const A xc = (A)x[idx];
  const int off[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  if (OPk == OP_UNIT) {
    A s = A(-6) * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + off[q][0], j + off[q][1], k + off[q][2]);
      bool neu = mn == BOUNDARY || (o.solid_c && mn == SOLID);
      A xn = (mn == FLUID || mn == AIR) ? (A)x[cid(g, i + off[q][0], j + off[q][1], k + off[q][2])] : A(0);
      s += neu ? xc : xn;
    }
    return s / ((A)g.h * (A)g.h);
  } else if (OPk == OP_VARDENS) {
    A s = A(0);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + off[q][0], j + off[q][1], k + off[q][2]);
      if (mn == BOUNDARY || (o.solid_c && mn == SOLID)) continue;
      int64_t nb = cid(g, i + off[q][0], j + off[q][1], k + off[q][2]);
      A xn = (mn == FLUID || mn == AIR) ? (A)x[nb] : A(0);
      s += face_beta<A>(o, idx, nb, m, mn) * (xn - xc);
    }
    return s / ((A)g.h * (A)g.h);
  } else {
    A s = ((A)o.hal[m] + A(6)) * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + off[q][0], j + off[q][1], k + off[q][2]);
      bool neu = mn == BOUNDARY || (o.vac_neumann && mn == VACUUM);
      s -= neu ? xc : (A)x[cid(g, i + off[q][0], j + off[q][1], k + off[q][2])];
    }
    return s;
  }
}

// Inverse Jacobi diagonal (1 for non-unknowns or when unpreconditioned).
template <class A, int OPk>
__device__ __forceinline__ A inv_diag(const OpArgs& o, int64_t idx, int precond) {
  const Grid& g = o.g;
  const int8_t m = o.mark[idx];
  if (!precond || !is_unknown(OPk, m)) return A(1);
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const int off[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  A s = A(0);
  if (OPk == OP_HEAT) {
    s = (A)o.hal[m] + A(6);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + off[q][0], j + off[q][1], k + off[q][2]);
      if (mn == BOUNDARY || (o.vac_neumann && mn == VACUUM)) s -= A(1);
    }
    return A(1) / s;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int8_t mn = mark_at(g, o.mark, i + off[q][0], j + off[q][1], k + off[q][2]);
    if (mn == BOUNDARY || (o.solid_c && mn == SOLID)) continue;
    s += (OPk == OP_UNIT) ? A(1) : face_beta<A>(o, idx, cid(g, i + off[q][0], j + off[q][1], k + off[q][2]), m, mn);
  }
  if (s == A(0)) return A(1);
  return -((A)g.h * (A)g.h) / s;
}

// ---- deterministic grid reduction with last-block finalisation ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double (*lds)[N]) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) lds[w][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = lds[0][q] + lds[1][q] + lds[2][q] + lds[3][q];
  __syncthreads();
}

// Writes this block's partial sums; returns true in every thread of the last
// block, with tot[] holding the grid totals (summed in block order).
template <int N>
__device__ bool grid_sum(double (&v)[N], double* __restrict__ partials, uint32_t* counter, double (&tot)[N]) {
  __shared__ double lds[PB / 64][N];
  __shared__ int is_last;
  block_sum<N>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) partials[(int64_t)blockIdx.x * N + q] = v[q];
    __threadfence();
    uint32_t old = atomicAdd(counter, 1u);
    is_last = (old == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double s[N];
#pragma unroll
  for (int q = 0; q < N; ++q) s[q] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += PB)
#pragma unroll
    for (int q = 0; q < N; ++q) s[q] += __builtin_nontemporal_load(&partials[(int64_t)b * N + q]);
  block_sum<N>(s, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) tot[q] = s[q];
    *counter = 0u;
  }
  return true;
}

__device__ __forceinline__ int64_t chunk_begin(const Launch& L) {
  const int nb = L.nblk;
  const int b = blockIdx.x;
  const int lb = (b & 7) * (nb >> 3) + (b >> 3);  // XCD-contiguous logical block
  return (int64_t)lb * L.chunk;
}

#define CELLS_OF_BLOCK(G)                                                        \
  const int64_t c0 = chunk_begin(L);                                             \
  const int64_t c1 = ((int64_t)(G) < c0 + L.chunk) ? (int64_t)(G) : c0 + L.chunk; \
  for (int64_t idx = c0 + threadIdx.x; idx < c1; idx += PB)

__device__ __forceinline__ bool bad(double v) { return !(v == v) || v == INFINITY || v == -INFINITY; }

// r = b (unknowns), x = 0, dinv, z = r*dinv, p = z; sums r.z, r.r, b.b.
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_init(OpArgs o, Launch L, const V* __restrict__ b, V* __restrict__ x,
                                             V* __restrict__ r, V* __restrict__ p, V* __restrict__ dinv,
                                             double* __restrict__ x64, int precond, double* partials,
                                             uint32_t* counter, SolveScalars* sc, int ref, double eps_abs,
                                             double tol) {
  Dacc a_rz = 0, a_rr = 0, a_bb = 0;
  CELLS_OF_BLOCK(o.g.G) {
    bool u = is_unknown(OPk, o.mark[idx]);
    V bv = u ? b[idx] : V(0);
    V di = inv_diag<V, OPk>(o, idx, precond);
    V z = bv * di;
    x[idx] = V(0);
    r[idx] = bv;
    p[idx] = z;
    dinv[idx] = di;
    if (x64) x64[idx] = 0.0;
    a_rz += (Dacc)bv * (Dacc)z;
    a_rr += (Dacc)bv * (Dacc)bv;
    a_bb += (Dacc)bv * (Dacc)bv;
  }
  double v[3] = {(double)a_rz, (double)a_rr, (double)a_bb}, t[3];
  if (grid_sum<3>(v, partials, counter, t) && threadIdx.x == 0) {
    sc->rz = t[0];
    sc->rr = t[1];
    sc->bb = t[2];
    sc->iter = 0;
    sc->refinements = 0;
    sc->refine_now = 0;
    sc->nonfinite = bad(t[0]) || bad(t[1]);
    bool conv = ref ? (t[1] < eps_abs) : (t[2] == 0.0 || sqrt(t[1]) <= tol * sqrt(t[2]));
    sc->converged = conv;
    sc->done = conv || sc->nonfinite;
  }
}

// q = A p ; p.q ; alpha = rz / pq
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_spmv(OpArgs o, Launch L, const V* __restrict__ p, V* __restrict__ q,
                                             double* partials, uint32_t* counter, SolveScalars* sc) {
  if (sc->done) return;
  Dacc a_pq = 0;
  CELLS_OF_BLOCK(o.g.G) {
    V qv = apply_op<V, OPk, V>(o, p, idx);
    q[idx] = qv;
    a_pq += (Dacc)p[idx] * (Dacc)qv;
  }
  double v[1] = {(double)a_pq}, t[1];
  if (grid_sum<1>(v, partials, counter, t) && threadIdx.x == 0) {
    sc->pq = t[0];
    if (t[0] == 0.0) {
      sc->done = 1;
      sc->converged = 1;
    } else if (bad(t[0])) {
      sc->done = 1;
      sc->nonfinite = 1;
    } else {
      sc->alpha = sc->rz / t[0];
    }
  }
}

// x += alpha p ; r -= alpha q ; z = r*dinv ; sums r.z, r.r ; convergence.
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_update(OpArgs o, Launch L, V* __restrict__ x, V* __restrict__ r,
                                               const V* __restrict__ p, const V* __restrict__ q,
                                               const V* __restrict__ dinv, double* partials, uint32_t* counter,
                                               SolveScalars* sc, int ref, double eps_abs, double tol, int cap,
                                               int mixed, int refine_every) {
  if (sc->done) return;
  const V al = (V)sc->alpha;
  Dacc a_rz = 0, a_rr = 0;
  CELLS_OF_BLOCK(o.g.G) {
    if (!is_unknown(OPk, o.mark[idx])) continue;
    V xv = x[idx] + al * p[idx];
    V rv = r[idx] - al * q[idx];
    x[idx] = xv;
    r[idx] = rv;
    V z = rv * dinv[idx];
    a_rz += (Dacc)rv * (Dacc)z;
    a_rr += (Dacc)rv * (Dacc)rv;
  }
  double v[2] = {(double)a_rz, (double)a_rr}, t[2];
  if (grid_sum<2>(v, partials, counter, t) && threadIdx.x == 0) {
    int it = sc->iter + 1;
    sc->iter = it;
    sc->rr = t[1];
    if (bad(t[0]) || bad(t[1])) {
      sc->done = 1;
      sc->nonfinite = 1;
      return;
    }
    bool conv = ref ? (t[1] < eps_abs) : (sqrt(t[1]) <= tol * sqrt(sc->bb));
    double rzp = sc->rz;
    sc->rz_new = rzp;  // rz_new holds the previous r.z here (reliable-update beta)
    sc->rz = t[0];
    sc->beta = t[0] / rzp;
    if (mixed) {
      // the FP64 true residual decides convergence and the cap
      if (conv || (refine_every > 0 && it % refine_every == 0) || it >= cap) sc->refine_now = 1;
    } else {
      if (conv) {
        sc->done = 1;
        sc->converged = 1;
      } else if (it >= cap) {
        sc->done = 1;
      }
    }
  }
}

// MIXED: x64 += x ; x = 0 (only when a refinement is due)
__global__ void __launch_bounds__(PB) k_fold(Launch L, int64_t G, float* __restrict__ x, double* __restrict__ x64,
                                             const SolveScalars* sc, int force) {
  if (!force && (sc->done || !sc->refine_now)) return;
  CELLS_OF_BLOCK(G) {
    x64[idx] += (double)x[idx];
    x[idx] = 0.0f;
  }
}

// MIXED: r = b - A x64 in FP64, stored to the FP32 residual; r.z and |r|^2 in
// FP64; convergence on the true residual; reliable-update beta.
template <int OPk>
__global__ void __launch_bounds__(PB) k_resid64(OpArgs o, Launch L, const float* __restrict__ b,
                                                const double* __restrict__ x64, float* __restrict__ r,
                                                const float* __restrict__ dinv, double* partials, uint32_t* counter,
                                                SolveScalars* sc, double tol, int cap) {
  if (sc->done || !sc->refine_now) return;
  double a_rz = 0, a_rr = 0;
  CELLS_OF_BLOCK(o.g.G) {
    bool u = is_unknown(OPk, o.mark[idx]);
    double rv = u ? ((double)b[idx] - apply_op<double, OPk, double>(o, x64, idx)) : 0.0;
    float rf = (float)rv;
    r[idx] = rf;
    a_rr += rv * rv;
    a_rz += (double)rf * (double)(rf * dinv[idx]);
  }
  double v[2] = {a_rz, a_rr}, t[2];
  if (grid_sum<2>(v, partials, counter, t) && threadIdx.x == 0) {
    sc->refine_now = 0;
    sc->refinements += 1;
    sc->rr_true = t[1];
    sc->rr = t[1];
    if (bad(t[0]) || bad(t[1])) {
      sc->done = 1;
      sc->nonfinite = 1;
      return;
    }
    if (sqrt(t[1]) <= tol * sqrt(sc->bb)) {
      sc->done = 1;
      sc->converged = 1;
    } else if (sc->iter >= cap) {
      sc->done = 1;
    } else {
      sc->rz = t[0];
      sc->beta = t[0] / sc->rz_new;
    }
  }
}

// p = r*dinv + beta p
template <class V>
__global__ void __launch_bounds__(PB) k_dir(Launch L, int64_t G, const V* __restrict__ r, const V* __restrict__ dinv,
                                            V* __restrict__ p, const SolveScalars* sc) {
  if (sc->done) return;
  const V be = (V)sc->beta;
  CELLS_OF_BLOCK(G) { p[idx] = r[idx] * dinv[idx] + be * p[idx]; }
}

// |b - A x|^2 in FP64 for the report.
template <int OPk, class B, class X>
__global__ void __launch_bounds__(PB) k_true_resid(OpArgs o, Launch L, const B* __restrict__ b, const X* __restrict__ x,
                                                   double* partials, uint32_t* counter, SolveScalars* sc) {
  double a_rr = 0;
  CELLS_OF_BLOCK(o.g.G) {
    bool u = is_unknown(OPk, o.mark[idx]);
    double rv = u ? ((double)b[idx] - apply_op<double, OPk, X>(o, x, idx)) : 0.0;
    a_rr += rv * rv;
  }
  double v[1] = {a_rr}, t[1];
  if (grid_sum<1>(v, partials, counter, t) && threadIdx.x == 0) sc->rr_true = t[0];
}

template <class V>
__global__ void k_copy_cast(int64_t n, const V* __restrict__ in, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

Launch make_launch(int64_t G) {
  Launch L;
  int64_t nb = (G + PB - 1) / PB;
  if (nb > MAX_BLK) nb = MAX_BLK;
  nb = ((nb + 7) / 8) * 8;
  L.nblk = (int)nb;
  int64_t ch = (G + nb - 1) / nb;
  L.chunk = ((ch + 63) / 64) * 64;
  return L;
}

OpArgs make_op(Ctx& c) {
  OpArgs o;
  o.g = c.g;
  o.mark = c.mark;
  o.ls = c.lsmerge;
  o.il = 1.0 / (double)c.cfg.rho_liquid;
  o.ia = 1.0 / (double)c.cfg.rho_air;
  const double h2 = (double)c.g.h * (double)c.g.h;
  for (int t = 0; t < 6; ++t)
    o.hal[t] = c.cfg.heat_alpha[t] > 0 ? h2 / (double)c.cfg.heat_alpha[t] / (double)c.cfg.dt : 0.0;
  o.solid_c = c.cfg.solid_constraint;
  o.vac_neumann = c.cfg.heat_vacuum_neumann;
  return o;
}

bool read_scalars(Ctx& c) {
  hipMemcpyAsync(c.scal_host, c.scal, sizeof(SolveScalars), hipMemcpyDeviceToHost, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "solver poll");
}

// Generic (P)CG driver.  V: vector type, Dacc: dot accumulation type.
// b: right-hand side (type V), x: solution (type V; MIXED adds x64).
template <class V, class Dacc, int OPk>
int cg_solve(Ctx& c, const V* b, V* x, double* x64, bool ref, int precond, int cap, double tol, bool mixed,
             mpf_solve_info* info) {
  const int64_t G = c.g.G;
  const Launch L = make_launch(G);
  const OpArgs o = make_op(c);
  V* r = (V*)c.vr;
  V* p = (V*)c.vp;
  V* q = (V*)c.vq;
  V* dinv = (V*)c.vdinv;
  uint32_t* ctr = c.counters;
  const double eps_abs = 1e-2 * (double)G;
  const int refine_every = c.cfg.refine_every;
  const int check = c.cfg.check_every > 0 ? c.cfg.check_every : 10;
  launch(c, "pcg_init", dim3(L.nblk), dim3(PB), 0, k_init<V, Dacc, OPk>, o, L, b, x, r, p, dinv,
         mixed ? x64 : (double*)nullptr, precond, c.partials, ctr, c.scal, (int)ref, eps_abs, tol);
  if (!read_scalars(c)) return MPF_HIP_ERROR;
  int launched = 0;
  while (!c.scal_host->done && launched < cap + 1) {
    int n = check;
    for (int s = 0; s < n; ++s) {
      launch(c, "pcg_spmv", dim3(L.nblk), dim3(PB), 0, k_spmv<V, Dacc, OPk>, o, L, (const V*)p, q, c.partials, ctr,
             c.scal);
      launch(c, "pcg_update", dim3(L.nblk), dim3(PB), 0, k_update<V, Dacc, OPk>, o, L, x, r, (const V*)p,
             (const V*)q, (const V*)dinv, c.partials, ctr, c.scal, (int)ref, eps_abs, tol, cap, (int)mixed,
             refine_every);
      if (mixed) {
        launch(c, "pcg_fold", dim3(L.nblk), dim3(PB), 0, k_fold, L, G, (float*)x, x64, (const SolveScalars*)c.scal, 0);
        launch(c, "pcg_resid64", dim3(L.nblk), dim3(PB), 0, k_resid64<OPk>, o, L, (const float*)b,
               (const double*)x64, (float*)r, (const float*)dinv, c.partials, ctr, c.scal, tol, cap);
      }
      launch(c, "pcg_dir", dim3(L.nblk), dim3(PB), 0, k_dir<V>, L, G, (const V*)r, (const V*)dinv, p,
             (const SolveScalars*)c.scal);
    }
    launched += n;
    if (!read_scalars(c)) return MPF_HIP_ERROR;
  }
  if (mixed) {
    launch(c, "pcg_fold", dim3(L.nblk), dim3(PB), 0, k_fold, L, G, (float*)x, x64, (const SolveScalars*)c.scal, 1);
    launch(c, "pcg_true_resid", dim3(L.nblk), dim3(PB), 0, k_true_resid<OPk, V, double>, o, L, b,
           (const double*)x64, c.partials, ctr, c.scal);
  } else {
    launch(c, "pcg_true_resid", dim3(L.nblk), dim3(PB), 0, k_true_resid<OPk, V, V>, o, L, b, (const V*)x, c.partials,
           ctr, c.scal);
  }
  if (!read_scalars(c)) return MPF_HIP_ERROR;
  const SolveScalars& s = *c.scal_host;
  if (info) {
    info->iterations = s.iter;
    info->refinements = s.refinements;
    info->converged = s.converged;
    info->rel_residual = s.bb > 0 ? sqrt(s.rr_true) / sqrt(s.bb) : 0.0;
    info->abs_rr = s.rr;
  }
  if (s.nonfinite) {
    c.err = "non-finite value in the solver";
    return MPF_NONFINITE;
  }
  return s.converged ? MPF_OK : MPF_NO_CONVERGENCE;
}

}  // namespace

int run_pressure_solve(Ctx& c, mpf_solve_info* info) {
  const mpf_config& k = c.cfg;
  const bool ref = k.solver == MPF_SOLVER_REFERENCE_CG;
  const int precond = ref ? 0 : 1;
  const int cap = ref ? k.ref_max_iters : k.max_iters;
  const bool vd = k.op == MPF_OPERATOR_VARDENS;
  if (k.precision == MPF_FP64) {
    if (vd) return cg_solve<double, double, OP_VARDENS>(c, (const double*)c.div, (double*)c.press, nullptr, ref, precond, cap, k.tol, false, info);
    return cg_solve<double, double, OP_UNIT>(c, (const double*)c.div, (double*)c.press, nullptr, ref, precond, cap, k.tol, false, info);
  }
  if (k.precision == MPF_MIXED) {
    if (vd) return cg_solve<float, double, OP_VARDENS>(c, (const float*)c.div, (float*)c.vx, (double*)c.press, ref, precond, cap, k.tol, true, info);
    return cg_solve<float, double, OP_UNIT>(c, (const float*)c.div, (float*)c.vx, (double*)c.press, ref, precond, cap, k.tol, true, info);
  }
  if (vd) return cg_solve<float, float, OP_VARDENS>(c, (const float*)c.div, (float*)c.press, nullptr, ref, precond, cap, k.tol, false, info);
  return cg_solve<float, float, OP_UNIT>(c, (const float*)c.div, (float*)c.press, nullptr, ref, precond, cap, k.tol, false, info);
}

// Heat: reference CG (no preconditioner, r.r < 1e-2 G), FP32 vectors, FP64 dots.
int run_heat_cg(Ctx& c, mpf_solve_info* info) {
  return cg_solve<float, double, OP_HEAT>(c, (const float*)c.ctmp, (float*)c.tp, nullptr, true, 0, c.cfg.heat_max_iters,
                                          0.0, false, info);
}

}  // namespace mpf
