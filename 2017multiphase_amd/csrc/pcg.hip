// Matrix-free (P)CG for the pressure projection (a12) and the implicit heat
// diffusion (a16).
//
// Replaces solver_cg_bubble (spray.cu:1613-1687) with computeAx_bubble /
// pcg_op_bubble / arrayproduct_k (spray_k.cu:3423-3460, 2279-2297) and
// solver_cg_heat (spray.cu:1780-1852) with computeAx_heat / pcg_op_heat
// (spray_k.cu:4526-4582).
//
// Two kernels per iteration, both HBM-bound:
//   spmv   : p_k = z_k + beta p_{k-1} (z = r * diag^-1), written to the other
//            p buffer; q = A p_k; p.q                       -> alpha
//            2.5-D tiling: a block owns a 16 x 64 (j, k) column tile over a
//            run of planes along i; each p_k is formed once, the current
//            plane (with a one-cell ring) sits in LDS for the j/k
//            neighbours and the i neighbours stay in registers.
//   update : x += alpha p_k ; r -= alpha q ; r.z ; r.r   -> beta, convergence
// FP32 vectors: spmv moves r 4 + p_{k-1} 4 + mask 1 + p_k 4 + q 4 = 17 B/cell,
// update x 8 + r 8 + p 4 + q 4 + mask 1 = 25 B/cell (FP64: 33 + 49).
// The Jacobi diagonal is rebuilt from a per-cell neighbour-mask byte (bits
// 0-5: face open, i.e. not Neumann; bit 6: unknown; bit 7: slow path — an air
// cell or a liquid/air face of the variable-density operator, which reads the
// flags and lsmerge) through a 7-entry table, so no diagonal vector is stored.
//
// Reductions are deterministic and reduced by the consumer: each kernel
// writes one partial per block with plain stores; every block of the next
// kernel sums all partials in the same fixed order and so derives the same
// alpha, beta and convergence decision.  The host polls the scalar record
// every check_every iterations.  MIXED adds, every refine_every iterations, a
// reliable update: the FP64 true residual b - A (x64 + x) replaces r (beta
// from it), and the next update folds the FP32 correction x into the FP64 x64.
// Streaming kernels walk contiguous cell chunks, mapped so each XCD (blocks
// are dispatched round-robin over the 8 XCDs) walks one contiguous slab.
#include "mpf_internal.h"

#include <type_traits>

namespace mpf {

namespace {

constexpr int PB = 512;        // threads per block
constexpr int MAX_BLK = 1024;  // blocks per solver launch
constexpr int MAX_RESTARTS = 50;  // MIXED: restarts from the FP64 true residual per solve
constexpr int OP_UNIT = 0, OP_VARDENS = 1, OP_HEAT = 2;
constexpr uint8_t NB_UNKNOWN = 64, NB_SLOW = 128;

struct OpArgs {
  Grid g;
  const int8_t* mark;
  const float* ls;
  const uint8_t* nbm;
  const void* dslow;  // 1/diag of slow rows (element type of the solve), filled with the mask
  double il, ia;    // 1/rho_liquid, 1/rho_air
  double hal[6];    // h^2/(alpha*dt) per cell type (heat)
  int solid_c;      // SOLID cells Neumann (pressure)
  int vac_neumann;  // VACUUM neighbours Neumann (heat)
  int precond;      // Jacobi preconditioner on
  // fast-path rows (uniform density): 1/diag for n open faces, and the factor
  // 1/(rho h^2) applied to the neighbour differences
  float dinv_f[7];
  double dinv_d[7];
  float qc_f;
  double qc_d;
};

template <class A> __device__ __forceinline__ A dinv_tab(const OpArgs& o, int n);
// (a chain of selects over register copies of the table: an indexed read of
// the kernel arguments would be a memory load, and waiting for it waits for
// every load issued before it; readfirstlane keeps the compiler from turning
// the selects back into a select of addresses)
__device__ __forceinline__ float uniform_reg(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double uniform_reg(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <class A>
__device__ __forceinline__ A select7(const A (&t)[7], int n) {
  A d = uniform_reg(t[0]);
#pragma unroll
  for (int c = 1; c < 7; ++c) d = (n == c) ? uniform_reg(t[c]) : d;
  return d;
}
template <> __device__ __forceinline__ float dinv_tab<float>(const OpArgs& o, int n) { return select7(o.dinv_f, n); }
template <> __device__ __forceinline__ double dinv_tab<double>(const OpArgs& o, int n) { return select7(o.dinv_d, n); }
template <class A> __device__ __forceinline__ A qcoef(const OpArgs& o);
template <> __device__ __forceinline__ float qcoef<float>(const OpArgs& o) { return o.qc_f; }
template <> __device__ __forceinline__ double qcoef<double>(const OpArgs& o) { return o.qc_d; }

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
__device__ __forceinline__ bool neumann(const OpArgs& o, int OPk, int8_t m) {
  if (OPk == OP_HEAT) return m == BOUNDARY || (o.vac_neumann && m == VACUUM);
  return m == BOUNDARY || (o.solid_c && m == SOLID);
}

__constant__ int c_off[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

// Neighbour-cell index offsets in the flat array, in the order of c_off.
__device__ __forceinline__ void nb_strides(const Grid& g, int64_t (&st)[6]) {
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  st[0] = sx; st[1] = sy; st[2] = 1; st[3] = -sx; st[4] = -sy; st[5] = -1;
}

// 1/rho of the face between cells ca (an unknown) and cb; a FLUID|AIR face
// uses theta of the lower-index cell so the operator is exactly symmetric.
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

// Neighbour mask of one cell.
// Generic operator row (flags and lsmerge), arithmetic A, x stored as X.
template <class A, int OPk, class X>
__device__ __forceinline__ A apply_generic(const OpArgs& o, const X* __restrict__ x, int64_t idx) {
  const Grid& g = o.g;
  const int8_t m = o.mark[idx];
  if (!is_unknown(OPk, m)) return A(0);
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const A xc = (A)x[idx];
  if (OPk == OP_HEAT) {
    A s = ((A)o.hal[m] + A(6)) * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
      s -= neumann(o, OPk, mn) ? xc : (A)x[cid(g, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2])];
    }
    return s;
  }
  A s = A(0);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int8_t mn = mark_at(g, o.mark, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
    if (neumann(o, OPk, mn)) continue;
    int64_t nb = cid(g, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
    A xn = (mn == FLUID || mn == AIR) ? (A)x[nb] : A(0);
    A be = (OPk == OP_UNIT) ? A(1) : face_beta<A>(o, idx, nb, m, mn);
    s += be * (xn - xc);
  }
  return s / ((A)g.h * (A)g.h);
}

template <class A, int OPk>
__device__ __forceinline__ A inv_diag_generic(const OpArgs& o, int64_t idx) {
  const Grid& g = o.g;
  const int8_t m = o.mark[idx];
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  A s = A(0);
  if (OPk == OP_HEAT) {
    s = (A)o.hal[m] + A(6);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int8_t mn = mark_at(g, o.mark, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
      if (neumann(o, OPk, mn)) s -= A(1);
    }
    return A(1) / s;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int8_t mn = mark_at(g, o.mark, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
    if (neumann(o, OPk, mn)) continue;
    s += (OPk == OP_UNIT) ? A(1)
                          : face_beta<A>(o, idx, cid(g, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]), m, mn);
  }
  if (s == A(0)) return A(1);
  return -((A)g.h * (A)g.h) / s;
}

template <int OPk, class V>
__global__ void k_nbm(OpArgs o, uint8_t* __restrict__ nbm, V* __restrict__ dslow) {
  const Grid& g = o.g;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const int8_t m = o.mark[idx];
  uint8_t b = 0;
  if (is_unknown(OPk, m)) b |= NB_UNKNOWN;
  bool slow = (OPk == OP_VARDENS) && m == AIR;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int8_t mn = mark_at(g, o.mark, i + c_off[q][0], j + c_off[q][1], k + c_off[q][2]);
    if (!neumann(o, OPk, mn)) b |= (uint8_t)(1u << q);
    if (OPk == OP_VARDENS && m == FLUID && mn == AIR) slow = true;
  }
  if (OPk == OP_HEAT) slow = true;  // the heat operator reads the flags (diffusivity per type)
  if (slow) b |= NB_SLOW;
  nbm[idx] = b;
  if (slow && (b & NB_UNKNOWN)) dslow[idx] = (V)inv_diag_generic<double, OPk>(o, idx);
}

// Slow-row diagonal as stored for a solve whose vectors are of type A.
template <class A>
__device__ __forceinline__ A dslow_at(const OpArgs& o, int64_t idx) {
  return ((const A*)o.dslow)[idx];
}

// 1/diag from the mask byte (fast) or the flags (slow); 1 off the unknowns or
// without preconditioning.
template <class A, int OPk>
__device__ __forceinline__ A inv_diag(const OpArgs& o, int64_t idx, uint8_t b) {
  if (!o.precond || !(b & NB_UNKNOWN)) return A(1);
  if (b & NB_SLOW) return dslow_at<A>(o, idx);
  return dinv_tab<A>(o, __popc((unsigned)(b & 63)));
}

// (A x)[idx] with the mask byte: fast path for liquid cells of uniform density.
template <class A, int OPk, class X>
__device__ __forceinline__ A apply_op(const OpArgs& o, const X* __restrict__ x, int64_t idx, uint8_t b) {
  if (!(b & NB_UNKNOWN)) return A(0);
  if (b & NB_SLOW) return apply_generic<A, OPk, X>(o, x, idx);
  int64_t st[6];
  nb_strides(o.g, st);
  const A xc = (A)x[idx];
  A s = A(0);
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if (b & (1u << q)) s += (A)x[idx + st[q]] - xc;
  return s * qcoef<A>(o);
}

// ---- deterministic grid reductions, reduced by the consumer ----
// A kernel writes one partial per block (plain stores, no fences or atomics);
// the next kernel's blocks each sum all partials in the same fixed order, so
// every block derives the same scalars (alpha, beta, convergence) bit for bit.
// Kernel boundaries make the partials visible across XCDs.  Block 0 also
// records the scalars for the host to poll.
// Wave total by DPP moves (no LDS round trip): a Hillis-Steele prefix sum
// within each row of 16 lanes (shifts 1, 2, 4, 8; lanes shifted in from
// outside the row read 0), then row 1 and 3 add the last lane of the row
// before them and rows 2 and 3 the last lane of row 1; lane 63 holds the
// total, which every lane reads.  The order is fixed, so every block that
// sums the same values gets the same bits.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xF, true);
  return v + __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0x111, 0xF>(v);  // row_shr:1
  v = dpp_add<0x112, 0xF>(v);  // row_shr:2
  v = dpp_add<0x114, 0xF>(v);  // row_shr:4
  v = dpp_add<0x118, 0xF>(v);  // row_shr:8
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, 63), hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Block-wide sum of N values over NT threads; the result is returned in every
// thread.  Thread q adds the wave sums of value q in wave order (one thread
// per value keeps the other threads' registers free of the NT/64 x N terms).
template <int N, int NT = PB>
__device__ __forceinline__ void block_sum_all(double (&v)[N]) {
  __shared__ double lds[NT / 64][N];
  __shared__ double tot[N];
  const int lane = __lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) lds[w][q] = v[q];
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < NT / 64; ++ww) s += lds[ww][threadIdx.x];
    tot[threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = tot[q];
  __syncthreads();
}

// This block's partial sums -> P[blockIdx.x * N + q].
template <int N, int NT = PB>
__device__ __forceinline__ void write_partials(double (&v)[N], double* __restrict__ P) {
  block_sum_all<N, NT>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) P[(int64_t)blockIdx.x * N + q] = v[q];
}

// Sum of the previous kernel's partials (nb blocks), identical in every block.
template <int N, int NT = PB>
__device__ __forceinline__ void read_partials(const double* __restrict__ P, int nb, double (&out)[N]) {
#pragma unroll
  for (int q = 0; q < N; ++q) out[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += NT)
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] += P[(int64_t)b * N + q];
  block_sum_all<N, NT>(out);
}

// Zeros for the (r.z, r.r) partials of blocks [nblk, npad) of a U slot, so
// that kernels launched with fewer blocks than npad leave npad summands.
template <int NT = PB>
__device__ __forceinline__ void pad_partials2(double* __restrict__ P, int nblk, int npad) {
  if (blockIdx.x == 0)
    for (int q = nblk + threadIdx.x; q < npad; q += NT) {
      P[2 * q] = 0.0;
      P[2 * q + 1] = 0.0;
    }
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

// Partial-sum buffers (device memory, doubles):
//   U[2][nblk][2]  (r.z, r.r) written by update(k) into slot k&1; init writes slot 1
//   S[nblk]        p.q written by spmv(k)
struct Parts {
  double* U0;
  double* U1;
  double* S;
  double* S2;  // p.q of the slow rows (k_slow_q)
};
__device__ __forceinline__ double* uslot(const Parts& P, int k) { return (k & 1) ? P.U1 : P.U0; }

// r = b (unknowns), x = 0 (and x64 = 0), the three p buffers 0 (p is stored
// on the unknowns only, and the slow-row kernels read it off them as 0);
// partials r.z, r.r into U slot 1.
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_init(OpArgs o, Launch L, const V* __restrict__ b, V* __restrict__ x,
                                             V* __restrict__ r, V* __restrict__ pa, V* __restrict__ pb,
                                             V* __restrict__ pc, double* __restrict__ x64, Parts P,
                                             SolveScalars* sc, int npad) {
  Dacc a_rz = 0, a_rr = 0;
  CELLS_OF_BLOCK(o.g.G) {
    const uint8_t nb = o.nbm[idx];
    V bv = (nb & NB_UNKNOWN) ? b[idx] : V(0);
    V z = bv * inv_diag<V, OPk>(o, idx, nb);
    x[idx] = V(0);
    r[idx] = bv;
    pa[idx] = V(0);
    pb[idx] = V(0);
    pc[idx] = V(0);
    if (x64) x64[idx] = 0.0;
    a_rz += (Dacc)bv * (Dacc)z;
    a_rr += (Dacc)bv * (Dacc)bv;
  }
  double v[2] = {(double)a_rz, (double)a_rr};
  write_partials<2>(v, uslot(P, 1));
  pad_partials2(uslot(P, 1), L.nblk, npad);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->done = 0;
    sc->converged = 0;
    sc->nonfinite = 0;
    sc->iter = 0;
    sc->refinements = 0;
    sc->last_true = -1;
  }
}

// ---- spmv tiling ----
constexpr int TJ = 8, TK = 128;               // (j, k) column tile
constexpr int ST = 512;                       // threads per spmv block
constexpr int RPT = TJ * TK / ST;             // tile rows per thread (2)
constexpr int LJ = TJ + 2, LK = TK + 2;       // LDS plane with a one-cell ring
constexpr int NRING = 2 * TK + 2 * TJ;        // ring cells per plane
static_assert(NRING <= ST, "one ring cell per thread");
static_assert((TK & (TK - 1)) == 0 && (ST / TK) * RPT == TJ, "threads cover the tile in rows of TK");

struct Tiling {
  int ntj, ntk, seg, nblk;
  int ns5;  // single-kernel iteration: blocks of slow-row partials after the nblk tile blocks
};
// Block -> tile run.  Blocks are dispatched round-robin over the 8 XCDs, so
// each XCD takes a contiguous run of tiles (its L2 sees neighbouring tiles);
// inside a run the second half swaps neighbouring tiles, so the two blocks a
// CU holds (run positions m and m + half) lie in different k tiles: a liquid
// surface across k makes k tiles unequal in work, and pairing a full with a
// partly filled tile evens the CUs' loads.  Any block order is correct.
__device__ __forceinline__ int tile_of_block(const Tiling& T) {
  const int b = blockIdx.x;
  if ((T.nblk & 31) == 0) {  // an even half, so the swap stays inside it
    int m = b >> 3;
    if (m >= (T.nblk >> 4)) m ^= 1;
    return (b & 7) * (T.nblk >> 3) + m;
  }
  if ((T.nblk & 7) == 0) return (b & 7) * (T.nblk >> 3) + (b >> 3);
  return b;
}


// The inputs of p_k at one cell: r, p_{k-1}, the slow-row diagonal and the
// mask byte (nb = 0 outside the grid).
template <class V>
struct Raw {
  V r, pp, ds;
  uint8_t nb;
};
// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a full __syncthreads would wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// a + x y in FP64 for FP32 x, y: their product is exact in FP64, so the fused
// form rounds exactly as a separate multiply and add (one instruction, not two).
template <class X, class Y>
__device__ __forceinline__ double acc64(double a, X x, Y y) {
  static_assert(std::is_same<X, float>::value && std::is_same<Y, float>::value, "FP32 factors only");
  return __builtin_fma((double)x, (double)y, a);
}

// Byte loads through a buffer resource: the array in scalar registers, the
// plane's offset as the scalar operand, the column in a vector register, so
// no lane forms a 64-bit address (mask arrays hold G < 2^32 bytes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t byte_rsrc(const uint8_t* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(p), 0, -1, 0x00020000);
}
__device__ __forceinline__ uint8_t ld_byte(__amdgpu_buffer_rsrc_t r, uint32_t col, uint32_t plane_off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, col, plane_off, 0);
}

// Vector stores that nothing reads again within the launch (the next launch
// or pass reads them): nontemporal, so they do not displace the lines that
// neighbouring tiles of the same launch are about to read.
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
template <class T>
__device__ __forceinline__ void st_nt4(T* p, const T (&e)[4]) {
  using T4 = T __attribute__((ext_vector_type(4)));
  const T4 t = {e[0], e[1], e[2], e[3]};
  __builtin_nontemporal_store(t, (T4*)p);
}

// q of a slow row (flags, per-face 1/rho, heat diagonal) from the p_k of its
// six neighbours (order +x, +y, +z, -x, -y, -z; 0 off the unknowns).
template <class V, int OPk>
__device__ __forceinline__ V slow_row(const OpArgs& o, int i, int j, int k, int64_t idx, V pcv, V n0, V n1, V n2, V n3,
                                      V n4, V n5) {
  const Grid& g = o.g;
  const V pn[6] = {n0, n1, n2, n3, n4, n5};
  const int8_t m = o.mark[idx];
  V s = V(0);
  if (OPk == OP_HEAT) s = ((V)o.hal[m] + V(6)) * pcv;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int8_t mn = mark_at(g, o.mark, i + c_off[f][0], j + c_off[f][1], k + c_off[f][2]);
    if (neumann(o, OPk, mn)) {
      if (OPk == OP_HEAT) s -= pcv;
      continue;
    }
    if (OPk == OP_HEAT) s -= pn[f];
    else {
      const int64_t n = cid(g, i + c_off[f][0], j + c_off[f][1], k + c_off[f][2]);
      const V bf = (OPk == OP_UNIT) ? V(1) : face_beta<V>(o, idx, n, m, mn);
      s += bf * (pn[f] - pcv);
    }
  }
  return (OPk == OP_HEAT) ? s : s / ((V)g.h * (V)g.h);
}

// Iteration k.  From update(k-1): rz_k, rr_k (slot (k-1)&1) and rz_{k-1}
// (slot k&1) -> convergence, beta.  Then p_k = r*dinv + beta p_{k-1} for the
// block's tile (plus its ring), stored; q = A p_k; partial p.q.
template <class V, class Dacc, int OPk, bool SLOW>
__global__ void __launch_bounds__(ST, (sizeof(V) == 4 && !SLOW) ? 6 : 4) k_spmv(OpArgs o, Tiling T, int nbu, const V* __restrict__ r,
                                             const V* __restrict__ pp, V* __restrict__ pc, V* __restrict__ q,
                                             Parts P, SolveScalars* sc, int k, int ref, double eps_abs, double tol,
                                             int cap) {
  if (sc->done) return;
  __shared__ V dtab[8];  // 1/diag of the fast rows by open-face count (ordered by read_partials' barriers)
  if (threadIdx.x < 8) dtab[threadIdx.x] = (threadIdx.x < 7) ? dinv_tab<V>(o, threadIdx.x) : V(1);
  double cur[2], prv[2];
  read_partials<2, ST>(uslot(P, k - 1), nbu, cur);
  if (k > 0) read_partials<2, ST>(uslot(P, k), nbu, prv);
  else prv[0] = prv[1] = 0.0;
  const double bb = (k == 0) ? cur[1] : sc->bb;
  const bool nonfin = bad(cur[0]) || bad(cur[1]);
  const bool conv = ref ? (cur[1] < eps_abs) : (cur[1] == 0.0 || sqrt(cur[1]) <= tol * sqrt(bb));
  if (conv || nonfin || k >= cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->converged = conv && !nonfin;
      sc->nonfinite = nonfin;
      sc->iter = k;
      sc->rr = cur[1];
      sc->rz = cur[0];
      if (k == 0) sc->bb = bb;
    }
    return;
  }
  const V be = (k == 0) ? V(0) : (V)(cur[0] / prv[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rz = cur[0];
    sc->rr = cur[1];
    sc->iter = k;
    if (k == 0) sc->bb = bb;
  }
  __shared__ V pl[3][LJ][LK];
  const Grid& g = o.g;
  const int lb = tile_of_block(T);
  const int tk = lb % T.ntk;
  const int rest = lb / T.ntk;
  const int tj = rest % T.ntj;
  const int i_lo = (rest / T.ntj) * T.seg;
  const int i_hi = min(i_lo + T.seg, g.nx);
  const int j0 = tj * TJ, k0 = tk * TK;
  const int tid = threadIdx.x;
  const int kl = tid & (TK - 1), jl0 = (tid / TK) * RPT;
  const int kk = k0 + kl;
  // this thread's ring cell (local coords in [-1, TJ] x [-1, TK])
  int hj = 0, hk = 0;
  const bool ring = tid < NRING;
  if (tid < 2 * TK) {
    hj = (tid < TK) ? -1 : TJ;
    hk = tid & (TK - 1);
  } else if (ring) {
    const int t2 = tid - 2 * TK;
    hj = t2 % TJ;
    hk = (t2 < TJ) ? -1 : TK;
  }
  // columns (offsets inside a plane, below 2^31) and whether they lie in the grid
  const int64_t sx = (int64_t)g.ny * g.nz;
  uint32_t col[RPT];
  bool cok[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int j = j0 + jl0 + u;
    cok[u] = j < g.ny && kk < g.nz;
    col[u] = cok[u] ? (uint32_t)(j * g.nz + kk) : 0u;
  }
  const bool hok = ring && (j0 + hj) >= 0 && (j0 + hj) < g.ny && (k0 + hk) >= 0 && (k0 + hk) < g.nz;
  const uint32_t hcol = hok ? (uint32_t)((j0 + hj) * g.nz + (k0 + hk)) : 0u;
  // Loads address a plane (a uniform base, clamped into the grid) plus a
  // 32-bit column offset.  Every lane issues every load, so the number of
  // loads in flight is the same on all paths: a lane with nothing to read
  // reads its plane's first element and discards it.  (The mask is kept as
  // loaded; load_v applies the grid test, so that no instruction touches the
  // byte before it is needed.)
  auto plane = [&](int i) -> int64_t { return (int64_t)min(max(i, 0), g.nx - 1) * sx; };
  const auto nbm_r = byte_rsrc(o.nbm);
  auto load_m = [&](int i, bool ok, uint32_t c) -> uint8_t {
    return ld_byte(nbm_r, (ok && i >= 0 && i < g.nx) ? c : 0u, (uint32_t)plane(i));
  };
  // vectors of one cell given its mask: r and p_{k-1} on the unknowns, the
  // diagonal on the slow unknowns
  auto load_v = [&](int i, bool ok, uint8_t raw, uint32_t c, Raw<V>& w) {
    const uint8_t nb = (ok && i >= 0 && i < g.nx) ? raw : (uint8_t)0;
    const int64_t pb = plane(i);
    const bool un = (nb & NB_UNKNOWN) != 0;
    const uint32_t cv = un ? c : 0u;
    w.nb = nb;
    w.r = (r + pb)[cv];
    w.pp = (pp + pb)[cv];
    // (a solve without slow rows reads no diagonals)
    w.ds = SLOW ? ((const V*)o.dslow + pb)[(un && (nb & NB_SLOW) && o.precond) ? c : 0u] : V(1);
  };
  auto form = [&](const Raw<V>& w) -> V {
    const V dt = dtab[__popc((unsigned)(w.nb & 63))];
    const V d = !o.precond ? V(1) : (w.nb & NB_SLOW) ? w.ds : dt;
    return (w.nb & NB_UNKNOWN) ? w.r * d + be * w.pp : V(0);
  };
  // Software pipeline over planes, three planes of loads ahead.  At step i:
  // the mask of plane i+4 and the vectors of plane i+3 (given its mask,
  // loaded one step earlier) are loaded; p of plane i+1 is formed from the
  // vectors loaded two steps earlier and written to LDS; q of plane i is
  // computed and p, q stored.  The barrier only orders LDS, so the loads of
  // planes i+2..i+4 stay in flight.  Per-plane registers live in three slots,
  // slot (plane - i_lo) % 3, and the loop is unrolled by a multiple of 3 so
  // that every slot index is a constant.
  constexpr int UNR = 6;
  Raw<V> R[3][RPT], HR[3];   // vectors of planes i+1 .. i+3
  uint8_t M[3][RPT], HM[3];  // masks of planes i+3, i+4
  V PV[3][RPT];              // p of planes i-1 .. i+1
  uint8_t NB[3][RPT];        // masks of planes i, i+1
  // prologue: p of planes i_lo-1 and i_lo, vectors of i_lo+1 and i_lo+2, mask
  // of i_lo+3; every mask first, then every vector load, then the arithmetic
  uint8_t MP[4][RPT], HP[3];  // masks of planes i_lo-1 .. i_lo+2
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
#pragma unroll
    for (int q = 0; q < 4; ++q) MP[q][u] = load_m(i_lo - 1 + q, cok[u], col[u]);
    M[0][u] = load_m(i_lo + 3, cok[u], col[u]);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) HP[q] = load_m(i_lo + q, hok, hcol);
  HM[0] = load_m(i_lo + 3, hok, hcol);
  Raw<V> W0[RPT], W1[RPT], WH;
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo - 1, cok[u], MP[0][u], col[u], W0[u]);
    load_v(i_lo, cok[u], MP[1][u], col[u], W1[u]);
  }
  load_v(i_lo, hok, HP[0], hcol, WH);
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo + 1, cok[u], MP[2][u], col[u], R[1][u]);
    load_v(i_lo + 2, cok[u], MP[3][u], col[u], R[2][u]);
  }
  load_v(i_lo + 1, hok, HP[1], hcol, HR[1]);
  load_v(i_lo + 2, hok, HP[2], hcol, HR[2]);
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    PV[2][u] = form(W0[u]);
    PV[0][u] = form(W1[u]);
    NB[0][u] = W1[u].nb;
    pl[i_lo % 3][jl0 + u + 1][kl + 1] = PV[0][u];
  }
  if (ring) pl[i_lo % 3][hj + 1][hk + 1] = form(WH);
  const V qc = qcoef<V>(o);
  Dacc a_pq = 0;
  for (int ib = i_lo; ib < i_hi; ib += UNR) {
#pragma unroll
    for (int d = 0; d < UNR; ++d) {
      const int i = ib + d;
      if (i < i_hi) {  // uniform over the block
        const int s0 = d % 3, s1 = (d + 1) % 3, s2 = (d + 2) % 3;  // slots of planes i, i+1, i+2 (i+3 = i)
        const int sn = (i + 1) % 3, scur = i % 3;
        // loads: mask of plane i+4 (slot s1), vectors of plane i+3 (slot s0)
#pragma unroll
        for (int u = 0; u < RPT; ++u) M[s1][u] = load_m(i + 4, cok[u], col[u]);
        HM[s1] = load_m(i + 4, hok, hcol);
#pragma unroll
        for (int u = 0; u < RPT; ++u) load_v(i + 3, cok[u], M[s0][u], col[u], R[s0][u]);
        load_v(i + 3, hok, HM[s0], hcol, HR[s0]);
        // form plane i+1
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          PV[s1][u] = form(R[s1][u]);
          NB[s1][u] = R[s1][u].nb;
          pl[sn][jl0 + u + 1][kl + 1] = PV[s1][u];
        }
        if (ring) pl[sn][hj + 1][hk + 1] = form(HR[s1]);
        lds_barrier();
        // q of plane i: qc times the sum over the open faces, in the order
        // +x, +y, +z, -x, -y, -z, of (p_n - p_c) (each difference exact
        // where p is smooth); slow rows: p only, q from k_slow_q
        const int64_t pb = (int64_t)i * sx;
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int jl = jl0 + u;
          const uint8_t nb = NB[s0][u];
          const V pcv = PV[s0][u];
          const V nyp = pl[scur][jl + 2][kl + 1], nzp = pl[scur][jl + 1][kl + 2];
          const V nym = pl[scur][jl][kl + 1], nzm = pl[scur][jl + 1][kl];
          V sum = V(0);
          sum = (nb & 1) ? sum + (PV[s1][u] - pcv) : sum;
          sum = (nb & 2) ? sum + (nyp - pcv) : sum;
          sum = (nb & 4) ? sum + (nzp - pcv) : sum;
          sum = (nb & 8) ? sum + (PV[s2][u] - pcv) : sum;
          sum = (nb & 16) ? sum + (nym - pcv) : sum;
          sum = (nb & 32) ? sum + (nzm - pcv) : sum;
          const bool fast = (nb & (NB_UNKNOWN | NB_SLOW)) == NB_UNKNOWN;
          const V qv = fast ? sum * qc : V(0);
          a_pq += (Dacc)pcv * (Dacc)qv;
          if (nb & NB_UNKNOWN) {
            st_nt(pc + pb + col[u], pcv);
            if (fast) st_nt(q + pb + col[u], qv);
          }
        }
      }
    }
  }
  double v[1] = {(double)a_pq};
  write_partials<1, ST>(v, P.S);
}

// Slow rows of the current solve as a list in cell order (flag, scan, place).
__global__ void k_slow_flag(int64_t G, const uint8_t* __restrict__ nbm, uint32_t* __restrict__ flag) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < G) flag[idx] = ((nbm[idx] & (NB_UNKNOWN | NB_SLOW)) == (NB_UNKNOWN | NB_SLOW)) ? 1u : 0u;
}
__global__ void k_slow_list(int64_t G, const uint8_t* __restrict__ nbm, const uint32_t* __restrict__ pos,
                            uint32_t* __restrict__ list) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < G && (nbm[idx] & (NB_UNKNOWN | NB_SLOW)) == (NB_UNKNOWN | NB_SLOW)) list[pos[idx]] = (uint32_t)idx;
}

// q = A p_k on the slow rows (after k_spmv has stored every p_k); partial p.q
// into S2.  A fixed grid walks the list, so the sums are deterministic.
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_slow_q(OpArgs o, const uint32_t* __restrict__ list, int64_t n,
                                               const V* __restrict__ pc, V* __restrict__ q, double* __restrict__ S2,
                                               SolveScalars* sc) {
  if (sc->done) return;
  const Grid& g = o.g;
  Dacc a_pq = 0;
  for (int64_t t = (int64_t)blockIdx.x * PB + threadIdx.x; t < n; t += (int64_t)gridDim.x * PB) {
    const int64_t idx = list[t];
    int i, j, k;
    ijk_of(g, idx, i, j, k);
    V pn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int ii = i + c_off[f][0], jj = j + c_off[f][1], kk = k + c_off[f][2];
      pn[f] = valid(g, ii, jj, kk) ? pc[cid(g, ii, jj, kk)] : V(0);
    }
    const V pcv = pc[idx];
    const V qv = slow_row<V, OPk>(o, i, j, k, idx, pcv, pn[0], pn[1], pn[2], pn[3], pn[4], pn[5]);
    q[idx] = qv;
    a_pq += (Dacc)pcv * (Dacc)qv;
  }
  double v[1] = {(double)a_pq};
  write_partials<1>(v, S2);
}

// x += alpha p_k (folding x into x64 right after a reliable update);
// r -= alpha q; partials r.z, r.r into U slot k&1.  x advances every other
// iteration: mode 0 leaves x (the host keeps alpha_k p_k pending), mode 1
// adds alpha_k p_k, mode 2 adds the pending alpha_{k-1} p_{k-1} and then
// alpha_k p_k, rounding after each add exactly like two updates.
template <class V, class Dacc, int OPk>
__global__ void __launch_bounds__(PB) k_update(OpArgs o, Launch L, V* __restrict__ x, V* __restrict__ r,
                                               const V* __restrict__ pc, const V* __restrict__ ppv,
                                               const V* __restrict__ q, double* __restrict__ x64, Parts P, int nbs,
                                               int nbs2, SolveScalars* sc, int k, int fold, int mode, int npad) {
  if (sc->done) return;
  double t[1], t2[1];
  read_partials<1>(P.S, nbs, t);
  read_partials<1>(P.S2, nbs2, t2);
  t[0] += t2[0];
  if (t[0] == 0.0 || bad(t[0])) {  // spray.cu:1654-1656: zp == 0 ends the solve
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->converged = t[0] == 0.0;
      sc->nonfinite = bad(t[0]);
      sc->iter = k;
    }
    return;
  }
  const double ald = sc->rz / t[0];
  const V al = (V)ald;
  const V al1 = (mode == 2) ? (V)sc->al[(k - 1) & 1] : V(0);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->al[k & 1] = ald;
  Dacc a_rz = 0, a_rr = 0;
  // one cell: x += al p (or x64 += x + al p, x = 0 when folding), after
  // x += al1 p_prev in mode 2; r -= al q; partial r.z and r.r
  auto cell = [&](uint8_t nb, V pv, V p1v, V xv, V rv, V qv, V ds, double x64v, V& xo, V& ro, double& x64o) {
    if (mode != 0) {
      const V xa = (mode == 2) ? xv + al1 * p1v : xv;
      xo = fold ? V(0) : xa + al * pv;
      if (fold) x64o = x64v + (double)(xa + al * pv);
    }
    const V rn = rv - al * qv;
    ro = rn;
    const V d = !o.precond ? V(1) : (nb & NB_SLOW) ? ds : dinv_tab<V>(o, __popc((unsigned)(nb & 63)));
    const V z = rn * d;
    a_rz += (Dacc)rn * (Dacc)z;
    a_rr += (Dacc)rn * (Dacc)rn;
  };
  // Quads of 4 consecutive cells per thread (one 16-byte access per FP32
  // vector, two per FP64 vector), UQ quads per pass.  The masks of the next
  // pass are loaded one pass ahead; the vectors of a quad are read only if it
  // holds an unknown, and a quad of unknowns is stored whole, others cell by
  // cell.
  constexpr int UQ = 2;
  struct alignas(4 * sizeof(V)) Q4 { V e[4]; };
  struct alignas(32) D4 { double e[4]; };
  const int64_t c0 = chunk_begin(L);
  const int64_t c1 = max(c0, (o.g.G < c0 + L.chunk) ? o.g.G : c0 + L.chunk);  // blocks past the grid: empty
  const int64_t cq = c0 + ((c1 - c0) & ~(int64_t)3);                            // end of the whole quads
  constexpr int64_t PASS = (int64_t)4 * UQ * PB;
  auto load_masks = [&](int64_t base, uint32_t (&m)[UQ]) {
#pragma unroll
    for (int u = 0; u < UQ; ++u) {
      const int64_t b4 = base + (int64_t)u * 4 * PB;
      m[u] = (b4 < cq) ? *(const uint32_t*)(o.nbm + b4) : 0u;
    }
  };
  uint32_t mn[UQ];
  load_masks(c0 + 4 * (int64_t)threadIdx.x, mn);
  for (int64_t base = c0 + 4 * (int64_t)threadIdx.x; base < cq; base += PASS) {
    uint32_t m4[UQ];
#pragma unroll
    for (int u = 0; u < UQ; ++u) m4[u] = mn[u];
    Q4 pv[UQ] = {}, p1[UQ] = {}, xv[UQ] = {}, rv[UQ], qv[UQ], ds[UQ];
    D4 x64v[UQ] = {};
#pragma unroll
    for (int u = 0; u < UQ; ++u) {
      const int64_t b4 = base + (int64_t)u * 4 * PB;
      if (m4[u] & 0x40404040u) {
        rv[u] = *(const Q4*)(r + b4);
        qv[u] = *(const Q4*)(q + b4);  // q is only defined on the unknowns; used only there
        if (mode != 0) {
          pv[u] = *(const Q4*)(pc + b4);
          xv[u] = *(const Q4*)(x + b4);
          if (mode == 2) p1[u] = *(const Q4*)(ppv + b4);
          if (fold) x64v[u] = *(const D4*)(x64 + b4);
        }
      }
    }
    load_masks(base + PASS, mn);
    // slow-row diagonals, for quads holding a slow row (the quad's other
    // cells read its first element)
#pragma unroll
    for (int u = 0; u < UQ; ++u) {
      const int64_t b4 = base + (int64_t)u * 4 * PB;
      if (o.precond && (m4[u] & 0x80808080u)) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint8_t nb = (uint8_t)(m4[u] >> (8 * e));
          const bool sl = (nb & (NB_UNKNOWN | NB_SLOW)) == (NB_UNKNOWN | NB_SLOW);
          ds[u].e[e] = dslow_at<V>(o, sl ? b4 + e : b4);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) ds[u].e[e] = V(1);
      }
    }
#pragma unroll
    for (int u = 0; u < UQ; ++u) {
      const uint32_t un4 = m4[u] & 0x40404040u;
      if (!un4) continue;
      const int64_t b4 = base + (int64_t)u * 4 * PB;
      Q4 xo, ro;
      D4 x64o;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint8_t nb = (uint8_t)(m4[u] >> (8 * e));
        xo.e[e] = xv[u].e[e];
        ro.e[e] = rv[u].e[e];
        x64o.e[e] = fold ? x64v[u].e[e] : 0.0;
        if (nb & NB_UNKNOWN)
          cell(nb, pv[u].e[e], p1[u].e[e], xv[u].e[e], rv[u].e[e], qv[u].e[e], ds[u].e[e], x64o.e[e], xo.e[e],
               ro.e[e], x64o.e[e]);
      }
      if (un4 == 0x40404040u) {
        st_nt4(r + b4, ro.e);
        if (mode != 0) st_nt4(x + b4, xo.e);
        if (fold) st_nt4(x64 + b4, x64o.e);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if ((m4[u] >> (8 * e)) & NB_UNKNOWN) {
            st_nt(r + b4 + e, ro.e[e]);
            if (mode != 0) st_nt(x + b4 + e, xo.e[e]);
            if (fold) st_nt(x64 + b4 + e, x64o.e[e]);
          }
      }
    }
  }
  // cells after the last whole quad (grids whose size is not a multiple of 4)
  if (threadIdx.x < c1 - cq) {
    const int64_t idx = cq + threadIdx.x;
    const uint8_t nb = o.nbm[idx];
    if (nb & NB_UNKNOWN) {
      const bool sl = o.precond && (nb & NB_SLOW);
      V xo = V(0), ro;
      double x64o = 0.0;
      cell(nb, mode != 0 ? pc[idx] : V(0), mode == 2 ? ppv[idx] : V(0), mode != 0 ? x[idx] : V(0), r[idx], q[idx],
           sl ? dslow_at<V>(o, idx) : V(1), fold ? x64[idx] : 0.0, xo, ro, x64o);
      r[idx] = ro;
      if (mode != 0) x[idx] = xo;
      if (fold) x64[idx] = x64o;
    }
  }
  double v[2] = {(double)a_rz, (double)a_rr};
  write_partials<2>(v, uslot(P, k));
  pad_partials2(uslot(P, k), L.nblk, npad);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->iter = k + 1;
}

// MIXED reliable update after update(k): r = b - A (x64 + x) in FP64, stored
// to the FP32 residual; (r.z, |r|^2) overwrite U slot k&1, so spmv(k+1)
// takes beta and the convergence test from the true residual.  The spmv's
// tiles and plane pipeline: x64 + x of each plane goes to LDS, fast rows sum
// their open faces from there, slow rows (flags and level set) also take
// their neighbours from LDS (0 off the unknowns).  Partials of npad blocks
// (zeros past the launch), the count the spmv reads.
//
// For the single-kernel iteration (q given): also q.z and q.r with the true
// residual, and the four sums (r.z, r.r, q.z, q.r) per block go to Pf.
// MODE 0: no slow rows; 1: slow rows inline (flags and level set read per
// row); 2: slow rows skipped (their residual and sums come from
// k_slow_resid64, whose partials follow this kernel's T.nblk blocks).
template <int OPk, int MODE>
__global__ void __launch_bounds__(ST, MODE == 1 ? 2 : 4) k_resid64(OpArgs o, Tiling T, const float* __restrict__ b,
                                                const double* __restrict__ x64, const float* __restrict__ x,
                                                float* __restrict__ r, Parts P, SolveScalars* sc, int k, int npad,
                                                const float* __restrict__ qf, double* __restrict__ Pf) {
  if (sc->done) return;
  __shared__ float dtab[8];
  if (threadIdx.x < 8) dtab[threadIdx.x] = (threadIdx.x < 7) ? dinv_tab<float>(o, threadIdx.x) : 1.0f;
  __shared__ double pl[3][LJ][LK];
  const Grid& g = o.g;
  const int lb = tile_of_block(T);
  const int tk = lb % T.ntk;
  const int rest = lb / T.ntk;
  const int tj = rest % T.ntj;
  const int i_lo = (rest / T.ntj) * T.seg;
  const int i_hi = min(i_lo + T.seg, g.nx);
  const int j0 = tj * TJ, k0 = tk * TK;
  const int tid = threadIdx.x;
  const int kl = tid & (TK - 1), jl0 = (tid / TK) * RPT;
  const int kk = k0 + kl;
  int hj = 0, hk = 0;
  const bool ring = tid < NRING;
  if (tid < 2 * TK) {
    hj = (tid < TK) ? -1 : TJ;
    hk = tid & (TK - 1);
  } else if (ring) {
    const int t2 = tid - 2 * TK;
    hj = t2 % TJ;
    hk = (t2 < TJ) ? -1 : TK;
  }
  const int64_t sx = (int64_t)g.ny * g.nz;
  uint32_t col[RPT];
  bool cok[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int j = j0 + jl0 + u;
    cok[u] = j < g.ny && kk < g.nz;
    col[u] = cok[u] ? (uint32_t)(j * g.nz + kk) : 0u;
  }
  const bool hok = ring && (j0 + hj) >= 0 && (j0 + hj) < g.ny && (k0 + hk) >= 0 && (k0 + hk) < g.nz;
  const uint32_t hcol = hok ? (uint32_t)((j0 + hj) * g.nz + (k0 + hk)) : 0u;
  // loads as in the spmv: uniform plane base, 32-bit column, every lane
  // issues every load (element 0 of the plane when there is nothing to read)
  auto plane = [&](int i) -> int64_t { return (int64_t)min(max(i, 0), g.nx - 1) * sx; };
  const auto nbm_r = byte_rsrc(o.nbm);
  auto load_m = [&](int i, bool ok, uint32_t c) -> uint8_t {
    return ld_byte(nbm_r, (ok && i >= 0 && i < g.nx) ? c : 0u, (uint32_t)plane(i));
  };
  struct Rw {
    double x64;
    float x, b;
    uint8_t nb;
  };
  auto load_v = [&](int i, bool ok, uint8_t raw, uint32_t c, Rw& w) {
    const uint8_t nb = (ok && i >= 0 && i < g.nx) ? raw : (uint8_t)0;
    const int64_t pb = plane(i);
    const uint32_t cv = (nb & NB_UNKNOWN) ? c : 0u;
    w.nb = nb;
    w.x64 = (x64 + pb)[cv];
    w.x = (x + pb)[cv];
    w.b = (b + pb)[cv];
  };
  auto form = [&](const Rw& w) -> double { return (w.nb & NB_UNKNOWN) ? w.x64 + (double)w.x : 0.0; };
  constexpr int UNR = 6;
  Rw R[3][RPT], HR[3];
  uint8_t M[3][RPT], HM[3];
  double XV[3][RPT];
  float BV[3][RPT];
  float QN[2][RPT];  // q of plane i (slot d&1), loaded one step ahead (single-kernel form only)
  uint8_t NB[3][RPT];
  // prologue: every mask of planes i_lo-1 .. i_lo+3 first, then every vector
  // load, then the arithmetic
  uint8_t MP[4][RPT], HP[3];  // masks of planes i_lo-1 .. i_lo+2
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
#pragma unroll
    for (int q = 0; q < 4; ++q) MP[q][u] = load_m(i_lo - 1 + q, cok[u], col[u]);
    M[0][u] = load_m(i_lo + 3, cok[u], col[u]);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) HP[q] = load_m(i_lo + q, hok, hcol);
  HM[0] = load_m(i_lo + 3, hok, hcol);
  Rw W0[RPT], W1[RPT], WH;
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo - 1, cok[u], MP[0][u], col[u], W0[u]);
    load_v(i_lo, cok[u], MP[1][u], col[u], W1[u]);
  }
  load_v(i_lo, hok, HP[0], hcol, WH);
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo + 1, cok[u], MP[2][u], col[u], R[1][u]);
    load_v(i_lo + 2, cok[u], MP[3][u], col[u], R[2][u]);
  }
  load_v(i_lo + 1, hok, HP[1], hcol, HR[1]);
  load_v(i_lo + 2, hok, HP[2], hcol, HR[2]);
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    XV[2][u] = form(W0[u]);
    XV[0][u] = form(W1[u]);
    NB[0][u] = W1[u].nb;
    BV[0][u] = W1[u].b;
    pl[i_lo % 3][jl0 + u + 1][kl + 1] = XV[0][u];
  }
  if (ring) pl[i_lo % 3][hj + 1][hk + 1] = form(WH);
  const double bc = (OPk == OP_UNIT) ? 1.0 : o.il;
  const double ih2 = 1.0 / ((double)g.h * (double)g.h);
  double a_rz = 0, a_rr = 0, a_qz = 0, a_qr = 0;
#pragma unroll
  for (int u = 0; u < RPT; ++u) QN[0][u] = qf ? (qf + plane(i_lo))[col[u]] : 0.0f;
  for (int ib = i_lo; ib < i_hi; ib += UNR) {
#pragma unroll
    for (int d = 0; d < UNR; ++d) {
      const int i = ib + d;
      if (i < i_hi) {  // uniform over the block
        const int s0 = d % 3, s1 = (d + 1) % 3, s2 = (d + 2) % 3;
        const int sn = (i + 1) % 3, scur = i % 3;
#pragma unroll
        for (int u = 0; u < RPT; ++u) M[s1][u] = load_m(i + 4, cok[u], col[u]);
        HM[s1] = load_m(i + 4, hok, hcol);
#pragma unroll
        for (int u = 0; u < RPT; ++u) load_v(i + 3, cok[u], M[s0][u], col[u], R[s0][u]);
        load_v(i + 3, hok, HM[s0], hcol, HR[s0]);
#pragma unroll
        for (int u = 0; u < RPT; ++u) QN[(d + 1) & 1][u] = qf ? (qf + plane(i + 1))[col[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          XV[s1][u] = form(R[s1][u]);
          NB[s1][u] = R[s1][u].nb;
          BV[s1][u] = R[s1][u].b;
          pl[sn][jl0 + u + 1][kl + 1] = XV[s1][u];
        }
        if (ring) pl[sn][hj + 1][hk + 1] = form(HR[s1]);
        lds_barrier();
        const int64_t pb = (int64_t)i * sx;
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int jl = jl0 + u;
          const uint8_t nb = NB[s0][u];
          if (!(nb & NB_UNKNOWN)) continue;
          if (MODE == 2 && (nb & NB_SLOW)) continue;
          const double xc = XV[s0][u];
          // neighbours in the order +x, +y, +z, -x, -y, -z
          const double xn[6] = {XV[s1][u], pl[scur][jl + 2][kl + 1], pl[scur][jl + 1][kl + 2],
                                XV[s2][u], pl[scur][jl][kl + 1], pl[scur][jl + 1][kl]};
          const int64_t idx = pb + col[u];
          double ax;
          if (MODE != 1 || !(nb & NB_SLOW)) {
            double sacc = 0.0;
#pragma unroll
            for (int f = 0; f < 6; ++f)
              if (nb & (1u << f)) sacc += xn[f] - xc;
            ax = (bc * sacc) * ih2;
          } else {
            int ci, cj, ck;
            ijk_of(g, idx, ci, cj, ck);
            const int8_t m = o.mark[idx];
            double sacc = 0.0;
#pragma unroll
            for (int f = 0; f < 6; ++f) {
              const int8_t mn = mark_at(g, o.mark, ci + c_off[f][0], cj + c_off[f][1], ck + c_off[f][2]);
              if (neumann(o, OPk, mn)) continue;
              const int64_t n = cid(g, ci + c_off[f][0], cj + c_off[f][1], ck + c_off[f][2]);
              const double bf = (OPk == OP_UNIT) ? 1.0 : face_beta<double>(o, idx, n, m, mn);
              sacc += bf * (xn[f] - xc);
            }
            ax = sacc / ((double)g.h * (double)g.h);
          }
          const double rv = (double)BV[s0][u] - ax;
          const float rf = (float)rv;
          st_nt(r + pb + col[u], rf);
          a_rr += rv * rv;
          const float dv = !o.precond ? 1.0f
                           : (MODE == 1 && (nb & NB_SLOW)) ? dslow_at<float>(o, idx)
                                                           : dtab[__popc((unsigned)(nb & 63))];
          a_rz = acc64(a_rz, rf, rf * dv);
          if (qf) {
            const float qv = QN[d & 1][u];
            a_qz = acc64(a_qz, qv, rf * dv);
            a_qr = acc64(a_qr, qv, rf);
          }
        }
      }
    }
  }
  if (qf) {
    double v[4] = {a_rz, a_rr, a_qz, a_qr};
    write_partials<4, ST>(v, Pf);
  } else {
    double v[2] = {a_rz, a_rr};
    write_partials<2, ST>(v, uslot(P, k));
    pad_partials2<ST>(uslot(P, k), T.nblk, npad);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->refinements += 1;
    sc->last_true = k;
  }
}

constexpr int FDOT = 7;  // r.z, r.r, p.q, q.z, q.Dq, q.r, q.q
// The scalars of one single-kernel iteration, identical in every block:
// alpha_{k-1}, beta_k and the stop decision from the previous launch's
// partial sums (see k_cgf); block 0 records them for the host.
struct CgScalars {
  double ald, be;
  bool stop, nonfin, zero_pq;
};
template <int NT>
__device__ __forceinline__ CgScalars cg_scalars(const Tiling& T, int nres, int first, int nbu, const double* __restrict__ Pprev,
                                                const double* __restrict__ Pres, SolveScalars* sc, int k, double tol,
                                                int cap) {
  double ald = 0.0, rz, rr, be = 0.0, rz_prev = 0.0;
  bool zero_pq = false, bad_pq = false;
  if (first == 1) {
    double v[2];
    read_partials<2, NT>(Pres, nbu, v);
    rz = v[0];
    rr = v[1];
  } else if (first == 2) {
    double v[4];
    read_partials<4, NT>(Pres, nres, v);
    rz = v[0];
    rr = v[1];
  } else {
    double d[FDOT];
    read_partials<FDOT, NT>(Pprev, T.nblk + T.ns5, d);
    if (Pres) {
      double e[4];
      read_partials<4, NT>(Pres, nres, e);
      d[0] = e[0];
      d[1] = e[1];
      d[3] = e[2];
      d[5] = e[3];
    }
    zero_pq = d[2] == 0.0;
    bad_pq = bad(d[2]);
    ald = zero_pq || bad_pq ? 0.0 : d[0] / d[2];
    rz_prev = d[0];
    rz = (d[0] - (2.0 * ald) * d[3]) + (ald * ald) * d[4];
    rr = (d[1] - (2.0 * ald) * d[5]) + (ald * ald) * d[6];
  }
  const double bb = (first == 1 && k == 0) ? rr : sc->bb;
  const bool nonfin = bad(rz) || bad(rr) || bad_pq;
  const double rrc = rr > 0.0 ? rr : 0.0;
  // an expanded r.z whose sign differs from the previous one (r.z has the
  // sign of the preconditioner for every nonzero r) has lost every
  // significant digit: the recurrence is spent, so the solve stops and the
  // host's FP64 true-residual check decides (it restarts from the true
  // residual with exact sums, or accepts x)
  // (also a step length that is not positive, or a recurrence residual grown
  // far past the right-hand side: conjugacy lost at the FP32 floor)
  const bool spent = !first && (!(rz * rz_prev > 0.0) || (!zero_pq && !(ald > 0.0)) || rr > 1e12 * bb);
  const bool conv = zero_pq || rrc == 0.0 || spent || sqrt(rrc) <= tol * sqrt(bb);
  const bool stop = conv || nonfin || k >= cap;
  if (!first && !stop) be = rz / rz_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rz = rz;
    sc->rr = rr;
    sc->iter = k;
    sc->alh[k & 1] = (float)ald;
    if (first == 1 && k == 0) sc->bb = bb;
    if (stop) {
      sc->done_k = k;
      __threadfence();  // done_k is visible before done
      sc->done = 1;
      sc->converged = conv && !nonfin;
      sc->nonfinite = nonfin;
    }
  }
  return CgScalars{ald, be, stop, nonfin, zero_pq};
}

// ---- single-kernel iteration (MIXED form, solves without slow rows) ----
// One launch per iteration k does what update(k-1) and spmv(k) do: with
// alpha_{k-1} = (r.z)_{k-1} / (p.q)_{k-1} it forms r_k = r_{k-1} - alpha q_{k-1},
// advances x by alpha p_{k-1}, and forms p_k = z_k + beta p_{k-1} and
// q_k = A p_k.  beta needs (r.z)_k before any r_k exists, so it is taken from
// the sums of the previous launch, which hold every term of
//   (r - a q).D(r - a q) = r.z - 2 a q.z + a^2 q.Dq   (and the same for r.r):
// each launch sums r.z, r.r of the r_k it formed (exact for its vectors) and
// p.q, q.z, q.Dq, q.r, q.q, so the expansion spans one iteration only.  r, q
// and p alternate between two buffers (ring cells read the previous values
// of a neighbouring tile).  first = 1: k = 0 (sums of the init kernel, U slot
// 1); first = 2: a restart from a true residual (sums of the residual pass).
// Pres (when not first): a true-residual pass ran after the previous launch;
// its r.z, r.r, q.z, q.r replace those of the previous launch.
//
// Solves with slow rows (SLOW): p_k of a slow row takes its own diagonal
// (loaded with the vectors); its q_k and its share of the q sums come from
// k_slow_q5 right after this launch, which writes them as T.ns5 more blocks
// after this launch's T.nblk partials, so the next launch sums all of them
// in one fixed order (likewise the true-residual sums of k_slow_resid64).
// STEADY: a launch with first = 0 and fold = 0 (all but one launch in
// refine_every), compiled with both as constants so the plane loop carries
// no branch on them.  XM (steady launches without slow rows): how x advances,
// 1: x += alpha_k-1 p_k-1; 2: not at all (that term stays pending, x is
// neither read nor written); 3: x = (x + alpha_k-2 p_k-2) + alpha_k-1 p_k-1,
// the pending term first (p2in = p_k-2, alpha_k-2 as launch k-1 applied it),
// so x takes the same roundings as when it advances every launch; 4: the
// launch after a true-residual pass (first = 0, fold = 1 as constants), whose
// x64 += x + alpha p reads x64 in the plane pipeline with the vectors.  A
// launch that stops the solve advances x by all it owes.
template <int OPk, bool SLOW, bool STEADY, int XM>
__global__ void __launch_bounds__(ST, 4) k_cgf(OpArgs o, Tiling T, int first, int nbu, const double* __restrict__ Pprev,
                                               const double* __restrict__ Pres, double* __restrict__ Pout,
                                               const float* __restrict__ rin, const float* __restrict__ qin,
                                               const float* __restrict__ pin, float* __restrict__ rout,
                                               float* __restrict__ qout, float* __restrict__ pout,
                                               float* __restrict__ x, double* __restrict__ x64, int fold,
                                               SolveScalars* sc, int k, double tol, int cap,
                                               const float* __restrict__ p2in) {
  static_assert(XM == 1 || (STEADY && !SLOW && (XM == 2 || XM == 3)) || (XM == 4 && !STEADY && !SLOW),
                "x modes 2 and 3 are steady launches, mode 4 the fold launch");
  using V = float;
  // a launch after the one that stopped the solve does nothing (the
  // stopping launch's own late blocks still finish its x update)
  if (sc->done && sc->done_k != k) return;
  if (STEADY) {
    first = 0;
    fold = 0;
  }
  if (XM == 4) {
    first = 0;
    fold = 1;
  }
  // 1/diag of the fast rows by open-face count (1 without a preconditioner)
  __shared__ V dtab[8];
  if (threadIdx.x < 8) dtab[threadIdx.x] = (threadIdx.x < 7 && o.precond) ? dinv_tab<V>(o, threadIdx.x) : V(1);
  const Grid& g = o.g;
  const int lb = tile_of_block(T);
  const int tk = lb % T.ntk;
  const int rest = lb / T.ntk;
  const int tj = rest % T.ntj;
  const int i_lo = (rest / T.ntj) * T.seg;
  const int i_hi = min(i_lo + T.seg, g.nx);
  const int j0 = tj * TJ, k0 = tk * TK;
  const int tid = threadIdx.x;
  const int kl = tid & (TK - 1), jl0 = (tid / TK) * RPT;
  const int kk = k0 + kl;
  const int64_t sx = (int64_t)g.ny * g.nz;
  uint32_t col[RPT];
  bool cok[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int j = j0 + jl0 + u;
    cok[u] = j < g.ny && kk < g.nz;
    col[u] = cok[u] ? (uint32_t)(j * g.nz + kk) : 0u;
  }
  auto plane = [&](int i) -> int64_t { return (int64_t)min(max(i, 0), g.nx - 1) * sx; };
  int hj = 0, hk = 0;
  const bool ring = tid < NRING;
  if (tid < 2 * TK) {
    hj = (tid < TK) ? -1 : TJ;
    hk = tid & (TK - 1);
  } else if (ring) {
    const int t2 = tid - 2 * TK;
    hj = t2 % TJ;
    hk = (t2 < TJ) ? -1 : TK;
  }
  const bool hok = ring && (j0 + hj) >= 0 && (j0 + hj) < g.ny && (k0 + hk) >= 0 && (k0 + hk) < g.nz;
  const uint32_t hcol = hok ? (uint32_t)((j0 + hj) * g.nz + (k0 + hk)) : 0u;
  const auto nbm_r = byte_rsrc(o.nbm);
  auto load_m = [&](int i, bool ok, uint32_t c) -> uint8_t {
    return ld_byte(nbm_r, (ok && i >= 0 && i < g.nx) ? c : 0u, (uint32_t)plane(i));
  };
  // masks of planes i_lo - 1 .. i_lo + 3 (the plane loop's slots M, HM and
  // the prologue's MA, HA, HB); they do not depend on alpha or beta
  uint8_t M[3][RPT], HM[3];
  uint8_t MA[2][RPT], HA, HB[2];
  auto load_masks = [&]() {
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      MA[0][u] = load_m(i_lo - 1, cok[u], col[u]);
      MA[1][u] = load_m(i_lo, cok[u], col[u]);
      M[1][u] = load_m(i_lo + 1, cok[u], col[u]);
      M[2][u] = load_m(i_lo + 2, cok[u], col[u]);
      M[0][u] = load_m(i_lo + 3, cok[u], col[u]);
    }
    HA = load_m(i_lo, hok, hcol);
    HB[0] = load_m(i_lo + 1, hok, hcol);
    HB[1] = load_m(i_lo + 2, hok, hcol);
    HM[0] = load_m(i_lo + 3, hok, hcol);
  };
  if constexpr (XM != 4) load_masks();
  const CgScalars cs = cg_scalars<ST>(T, T.nblk + T.ns5, first, nbu, Pprev, Pres, sc, k, tol, cap);
  const bool stop = cs.stop, nonfin = cs.nonfin, zero_pq = cs.zero_pq;
  const V al = (V)cs.ald, bev = (V)cs.be;
  const V al2 = (XM == 3) ? sc->alh[(k - 1) & 1] : V(0);  // alpha_k-2, as launch k-1 applied it
  __shared__ V pl[3][LJ][LK];
  // x += alpha p_{k-1} on the unknowns (x64 += x + alpha p, x = 0 when folding)
  auto advance_x = [&](int64_t idx, V pv) {
    const V xv = x[idx];
    if (fold) {
      x64[idx] = x64[idx] + (double)(xv + al * pv);
      x[idx] = V(0);
    } else if (!first) {
      if (XM == 3) x[idx] = (xv + al2 * p2in[idx]) + al * pv;
      else x[idx] = xv + al * pv;  // a stopping launch advances x in mode 2 as well
    }
  };
  if (stop) {
    if (nonfin || zero_pq) return;
    for (int i = i_lo; i < i_hi; ++i)
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        if (!cok[u]) continue;
        const int64_t idx = (int64_t)i * sx + col[u];
        if (o.nbm[idx] & NB_UNKNOWN) advance_x(idx, pin[idx]);
      }
    return;
  }
  // loads as in the spmv: uniform plane base, 32-bit column, every lane
  // issues every load (element 0 of the plane when there is nothing to read)
  // (the slow-row diagonal exists only in the SLOW form: a field that is
  // constant there still shifts the register allocation of the fast form)
  struct RwS {
    V r, q, p, x, ds;
    uint8_t nb;
  };
  struct RwF {
    V r, q, p, x, p2;
    uint8_t nb;
  };
  using Rw = typename std::conditional<SLOW, RwS, RwF>::type;
  auto load_v = [&](int i, bool ok, uint8_t raw, uint32_t c, Rw& w, bool with_x) {
    const uint8_t nb = (ok && i >= 0 && i < g.nx) ? raw : (uint8_t)0;
    const int64_t pb = plane(i);
    const bool un = (nb & NB_UNKNOWN) != 0;
    const uint32_t cv = un ? c : 0u;
    w.nb = nb;
    w.r = (rin + pb)[cv];
    w.q = (qin + pb)[cv];
    w.p = (pin + pb)[cv];
    w.x = (with_x && XM != 2) ? (x + pb)[cv] : V(0);
    if constexpr (SLOW) w.ds = ((const V*)o.dslow + pb)[(un && (nb & NB_SLOW) && o.precond) ? c : 0u];
    else w.p2 = (with_x && XM == 3) ? (p2in + pb)[cv] : V(0);
  };
  // r_k, and p_k = z_k + beta p_{k-1} (0 off the unknowns)
  auto r_of = [&](const Rw& w) -> V { return first ? w.r : w.r - al * w.q; };
  auto d_of = [&](uint8_t nb) -> V { return dtab[__popc((unsigned)(nb & 63))]; };
  auto dw_of = [&](const Rw& w) -> V {
    if constexpr (SLOW) {
      if (w.nb & NB_SLOW) return o.precond ? w.ds : V(1);
    }
    return d_of(w.nb);
  };
  auto p_of = [&](const Rw& w, V rk) -> V {
    if (!(w.nb & NB_UNKNOWN)) return V(0);
    return first ? rk * dw_of(w) : rk * dw_of(w) + bev * w.p;
  };
  constexpr int UNR = 6;
  static_assert(UNR % 6 == 0, "slots by d % 3 and d & 1 repeat across the unrolled loop");
  Rw R[3][RPT], HR[3];
  V PV[3][RPT], RV[3][RPT];  // p_k, r_k of planes i-1 .. i+1
  uint8_t NB[3][RPT];
  double a[FDOT] = {0, 0, 0, 0, 0, 0, 0};
  // x64 of planes i+1 and i+2 in mode 4 (slot by plane parity within the
  // unrolled loop), loaded a plane ahead of their use
  double XD[2][RPT] = {};
  // a cell of this block's planes once its r_k is formed: r_k and x stored,
  // r.z and r.r summed (xd: its x64 in mode 4)
  auto own = [&](int i, int u, const Rw& w, V rk, double xd) {
    if (!(w.nb & NB_UNKNOWN)) return;
    const int64_t pb = (int64_t)i * sx;
    st_nt(rout + pb + col[u], rk);
    if (fold) {
      const int64_t idx = pb + col[u];
      if constexpr (XM == 4) st_nt(x64 + idx, xd + (double)(w.x + al * w.p));
      else x64[idx] = x64[idx] + (double)(w.x + al * w.p);
      st_nt(x + idx, V(0));
    } else if (!first) {
      if constexpr (XM == 1) st_nt(x + pb + col[u], w.x + al * w.p);
      if constexpr (XM == 3) st_nt(x + pb + col[u], (w.x + al2 * w.p2) + al * w.p);
    }
    const V zv = rk * dw_of(w);
    a[0] = acc64(a[0], rk, zv);
    a[1] = acc64(a[1], rk, rk);
  };
  // prologue: the masks of planes i_lo - 1 .. i_lo + 3 first (before the
  // partial sums are read, except in the fold launch, whose registers are
  // full), then every vector load, then the arithmetic
  if constexpr (XM == 4) load_masks();
  Rw W0[RPT], W1[RPT], WH;
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo - 1, cok[u], MA[0][u], col[u], W0[u], false);
    load_v(i_lo, cok[u], MA[1][u], col[u], W1[u], true);
  }
  load_v(i_lo, hok, HA, hcol, WH, false);
  double X0[RPT] = {};
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    load_v(i_lo + 1, cok[u], M[1][u], col[u], R[1][u], true);
    load_v(i_lo + 2, cok[u], M[2][u], col[u], R[2][u], true);
    if constexpr (XM == 4) {
      const uint8_t nb0 = (cok[u] && i_lo < g.nx) ? MA[1][u] : (uint8_t)0;
      X0[u] = (x64 + plane(i_lo))[(nb0 & NB_UNKNOWN) ? col[u] : 0u];
      const uint8_t nb1 = (cok[u] && i_lo + 1 < g.nx) ? M[1][u] : (uint8_t)0;
      XD[1][u] = (x64 + plane(i_lo + 1))[(nb1 & NB_UNKNOWN) ? col[u] : 0u];
    }
  }
  load_v(i_lo + 1, hok, HB[0], hcol, HR[1], false);
  load_v(i_lo + 2, hok, HB[1], hcol, HR[2], false);
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    PV[2][u] = p_of(W0[u], r_of(W0[u]));
    RV[0][u] = r_of(W1[u]);
    PV[0][u] = p_of(W1[u], RV[0][u]);
    NB[0][u] = W1[u].nb;
    pl[i_lo % 3][jl0 + u + 1][kl + 1] = PV[0][u];
    if (i_lo < i_hi) own(i_lo, u, W1[u], RV[0][u], X0[u]);
  }
  if (ring) pl[i_lo % 3][hj + 1][hk + 1] = p_of(WH, r_of(WH));
  const V qc = qcoef<V>(o);
  for (int ib = i_lo; ib < i_hi; ib += UNR) {
#pragma unroll
    for (int d = 0; d < UNR; ++d) {
      const int i = ib + d;
      if (i < i_hi) {  // uniform over the block
        const int s0 = d % 3, s1 = (d + 1) % 3, s2 = (d + 2) % 3;
        const int sn = (i + 1) % 3, scur = i % 3;
#pragma unroll
        for (int u = 0; u < RPT; ++u) M[s1][u] = load_m(i + 4, cok[u], col[u]);
        HM[s1] = load_m(i + 4, hok, hcol);
#pragma unroll
        for (int u = 0; u < RPT; ++u) load_v(i + 3, cok[u], M[s0][u], col[u], R[s0][u], true);
        load_v(i + 3, hok, HM[s0], hcol, HR[s0], false);
        if constexpr (XM == 4) {  // x64 of plane i+2 (its mask M[s2] was read two planes ago)
#pragma unroll
          for (int u = 0; u < RPT; ++u) {
            const uint8_t nb2 = (cok[u] && i + 2 < g.nx) ? M[s2][u] : (uint8_t)0;
            XD[d & 1][u] = (x64 + plane(i + 2))[(nb2 & NB_UNKNOWN) ? col[u] : 0u];
          }
        }
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          RV[s1][u] = r_of(R[s1][u]);
          PV[s1][u] = p_of(R[s1][u], RV[s1][u]);
          NB[s1][u] = R[s1][u].nb;
          pl[sn][jl0 + u + 1][kl + 1] = PV[s1][u];
          if (i + 1 < i_hi) own(i + 1, u, R[s1][u], RV[s1][u], XD[(d + 1) & 1][u]);
        }
        if (ring) pl[sn][hj + 1][hk + 1] = p_of(HR[s1], r_of(HR[s1]));
        lds_barrier();
        const int64_t pb = (int64_t)i * sx;
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int jl = jl0 + u;
          const uint8_t nb = NB[s0][u];
          const V pcv = PV[s0][u];
          const V nyp = pl[scur][jl + 2][kl + 1], nzp = pl[scur][jl + 1][kl + 2];
          const V nym = pl[scur][jl][kl + 1], nzm = pl[scur][jl + 1][kl];
          V sum = V(0);
          sum = (nb & 1) ? sum + (PV[s1][u] - pcv) : sum;
          sum = (nb & 2) ? sum + (nyp - pcv) : sum;
          sum = (nb & 4) ? sum + (nzp - pcv) : sum;
          sum = (nb & 8) ? sum + (PV[s2][u] - pcv) : sum;
          sum = (nb & 16) ? sum + (nym - pcv) : sum;
          sum = (nb & 32) ? sum + (nzm - pcv) : sum;
          if (SLOW && (nb & NB_SLOW)) {
            // slow row: p_k only, q_k from k_slow_q5
            if (nb & NB_UNKNOWN) st_nt(pout + pb + col[u], pcv);
          } else if (nb & NB_UNKNOWN) {
            const V qv = sum * qc;
            const V rk = RV[s0][u];
            const V dv = d_of(nb);
            const V zv = rk * dv;
            st_nt(pout + pb + col[u], pcv);
            st_nt(qout + pb + col[u], qv);
            a[2] = acc64(a[2], pcv, qv);
            a[3] = acc64(a[3], qv, zv);
            a[4] = acc64(a[4], qv, dv * qv);
            a[5] = acc64(a[5], qv, rk);
            a[6] = acc64(a[6], qv, qv);
          }
        }
      }
    }
  }
  write_partials<FDOT, ST>(a, Pout);
}

// ---- single-kernel iteration without a q vector (MIXED form, fast rows) ----
// The iteration of k_cgf from r_{k-1}, p_{k-1} and x alone: q_{k-1} = A p_{k-1}
// is formed again, on the tile and its first ring, from a p_{k-1} plane with
// two rings in LDS, instead of being read; q_k is stored only when the caller
// passes qout (the launch before a true-residual pass, which sums q.z and
// q.r).  Per unknown it moves r (read and written), p_{k-1} (read), p_k
// (written) and x (read and written): 24 bytes where k_cgf moves 32.  Every
// cell value is the same expression on the same operands as in k_cgf (q_{k-1}
// is the stencil k_cgf evaluated for q_k one launch earlier); the dot products
// group their terms by other blocks, so the scalars, and with them the
// iterates, agree with k_cgf's to rounding.
//
// One 1024-thread block per tile run (one block per CU): a thread holds one
// tile cell and at most one ring cell of the first or the second ring, so the
// longer pipeline fits the register file at four waves per SIMD.  Per plane i
// of the run, behind one LDS barrier: the p_{k-1} plane i+2 goes to LDS (tile
// and both rings); r_k and p_k of plane i+1 are formed on the tile and the
// first ring (q_{k-1} from the LDS plane and the x neighbours held in
// registers) and p_k goes to LDS; q_k of plane i is formed on the tile from
// the p_k plane.  Loads run two planes ahead of their use (masks three).
constexpr int SQ = 1024;                     // threads per block
constexpr int LJ2 = TJ + 4, LK2 = TK + 4;    // LDS plane with two rings
constexpr int NRING2 = 2 * TK + 2 * TJ + 4;  // second ring + the first ring's corners
static_assert(TJ * TK == SQ, "one tile cell per thread");
static_assert(NRING + NRING2 <= SQ, "at most one ring cell per thread");

template <int OPk>
__global__ void __launch_bounds__(SQ, 4) k_cgq(OpArgs o, Tiling T, int nres, int first, int nbu, const double* __restrict__ Pprev,
                                            const double* __restrict__ Pres, double* __restrict__ Pout,
                                            const float* __restrict__ rin, const float* __restrict__ pin,
                                            float* __restrict__ rout, float* __restrict__ pout,
                                            float* __restrict__ qout, float* __restrict__ x,
                                            double* __restrict__ x64, int fold, SolveScalars* sc, int k, double tol,
                                            int cap) {
  using V = float;
  if (sc->done && sc->done_k != k) return;
  __shared__ V dtab[8];
  if (threadIdx.x < 8) dtab[threadIdx.x] = (threadIdx.x < 7) ? dinv_tab<V>(o, threadIdx.x) : V(1);
  const CgScalars cs = cg_scalars<SQ>(T, nres, first, nbu, Pprev, Pres, sc, k, tol, cap);
  const V al = (V)cs.ald, bev = (V)cs.be;
  __shared__ V pa[2][LJ2][LK2];  // p_{k-1}: planes with two rings
  __shared__ V pk[2][LJ][LK];    // p_k: planes with one ring
  const Grid& g = o.g;
  const int lb = tile_of_block(T);
  const int tk = lb % T.ntk;
  const int rest = lb / T.ntk;
  const int tj = rest % T.ntj;
  const int i_lo = (rest / T.ntj) * T.seg;
  const int i_hi = min(i_lo + T.seg, g.nx);
  const int j0 = tj * TJ, k0 = tk * TK;
  const int tid = threadIdx.x;
  const int kl = tid & (TK - 1), jl = tid / TK;
  const int64_t sx = (int64_t)g.ny * g.nz;
  const bool cok = (j0 + jl) < g.ny && (k0 + kl) < g.nz;
  const uint32_t col = cok ? (uint32_t)((j0 + jl) * g.nz + k0 + kl) : 0u;
  auto plane = [&](int i) __attribute__((always_inline)) -> int64_t { return (int64_t)min(max(i, 0), g.nx - 1) * sx; };
  if (cs.stop) {
    if (cs.nonfin || cs.zero_pq || !cok) return;
    for (int i = i_lo; i < i_hi; ++i) {
      const int64_t idx = (int64_t)i * sx + col;
      if (o.nbm[idx] & NB_UNKNOWN) {
        const V xv = x[idx];
        if (fold) {
          x64[idx] = x64[idx] + (double)(xv + al * pin[idx]);
          x[idx] = V(0);
        } else if (!first) {
          x[idx] = xv + al * pin[idx];
        }
      }
    }
    return;
  }
  // the thread's ring cell (hj, hk): first ring for tid < NRING (as in
  // k_cgf), second ring and the first ring's corners for the next NRING2
  int hj = 0, hk = 0;
  const bool r1 = tid < NRING;
  const bool r2 = !r1 && tid < NRING + NRING2;
  if (tid < 2 * TK) {
    hj = (tid < TK) ? -1 : TJ;
    hk = tid & (TK - 1);
  } else if (r1) {
    const int t2 = tid - 2 * TK;
    hj = t2 % TJ;
    hk = (t2 < TJ) ? -1 : TK;
  } else if (r2) {
    const int t = tid - NRING;
    if (t < 2 * TK) {
      hj = (t < TK) ? -2 : TJ + 1;
      hk = t & (TK - 1);
    } else if (t < 2 * TK + 2 * TJ) {
      const int t2 = t - 2 * TK;
      hj = t2 % TJ;
      hk = (t2 < TJ) ? -2 : TK + 1;
    } else {
      const int t3 = t - 2 * TK - 2 * TJ;
      hj = (t3 & 1) ? TJ : -1;
      hk = (t3 & 2) ? TK : -1;
    }
  }
  const bool hr = r1 || r2;
  const bool hok = hr && (j0 + hj) >= 0 && (j0 + hj) < g.ny && (k0 + hk) >= 0 && (k0 + hk) < g.nz;
  const uint32_t hcol = hok ? (uint32_t)((j0 + hj) * g.nz + (k0 + hk)) : 0u;
  // loads as in k_cgf: uniform plane base, 32-bit column, every lane issues
  // every load (element 0 of the plane off the unknowns; r and x only where
  // the cell forms r_k)
  struct W {
    V r, p, x;
    uint8_t nb;
  };
  auto load_m = [&](int i, bool ok, uint32_t c) __attribute__((always_inline)) -> uint8_t {
    const uint8_t* mp = o.nbm + plane(i);
    return mp[(ok && i >= 0 && i < g.nx) ? c : 0u];
  };
  auto load_v = [&](int i, bool ok, uint8_t raw, uint32_t c, W& w, bool with_r, bool with_x) __attribute__((always_inline)) {
    const uint8_t nb = (ok && i >= 0 && i < g.nx) ? raw : (uint8_t)0;
    const int64_t pb = plane(i);
    const uint32_t cv = (nb & NB_UNKNOWN) ? c : 0u;
    w.nb = nb;
    w.p = (pin + pb)[cv];
    w.r = (rin + pb)[with_r ? cv : 0u];
    w.x = with_x ? (x + pb)[cv] : V(0);
  };
  auto pval = [&](const W& w) __attribute__((always_inline)) -> V { return (w.nb & NB_UNKNOWN) ? w.p : V(0); };
  auto d_of = [&](uint8_t nb) __attribute__((always_inline)) -> V { return o.precond ? dtab[__popc((unsigned)(nb & 63))] : V(1); };
  const V qc = qcoef<V>(o);
  // A p at a cell: neighbours +x, +y, +z, -x, -y, -z over its open faces, in
  // k_cgf's order and operations
  auto stencil = [&](uint8_t nb, V pc, V xp, V yp, V zp, V xm, V ym, V zm) __attribute__((always_inline)) -> V {
    V sum = V(0);
    sum = (nb & 1) ? sum + (xp - pc) : sum;
    sum = (nb & 2) ? sum + (yp - pc) : sum;
    sum = (nb & 4) ? sum + (zp - pc) : sum;
    sum = (nb & 8) ? sum + (xm - pc) : sum;
    sum = (nb & 16) ? sum + (ym - pc) : sum;
    sum = (nb & 32) ? sum + (zm - pc) : sum;
    return sum * qc;
  };
  // r_k of a cell from its loads and q_{k-1} (first: r_k = r_{k-1}), p_k
  auto r_new = [&](const W& w, V qprev) __attribute__((always_inline)) -> V { return first ? w.r : w.r - al * qprev; };
  auto p_new = [&](const W& w, V rk) __attribute__((always_inline)) -> V {
    if (!(w.nb & NB_UNKNOWN)) return V(0);
    return first ? rk * d_of(w.nb) : rk * d_of(w.nb) + bev * w.p;
  };
  double a[FDOT] = {0, 0, 0, 0, 0, 0, 0};
  auto own = [&](int i, const W& w, V rk) __attribute__((always_inline)) {
    if (!(w.nb & NB_UNKNOWN)) return;
    const int64_t pb = (int64_t)i * sx;
    st_nt(rout + pb + col, rk);
    if (fold) {
      const int64_t idx = pb + col;
      x64[idx] = x64[idx] + (double)(w.x + al * w.p);
      (x + pb)[col] = V(0);
    } else if (!first) {
      st_nt(x + pb + col, w.x + al * w.p);
    }
    const V zv = rk * d_of(w.nb);
    a[0] = acc64(a[0], rk, zv);
    a[1] = acc64(a[1], rk, rk);
  };
  // registers by plane q in slot (q - i_lo) & 3: loads L (tile cell), H
  // (ring cell); p_{k-1} PP, PH; p_k, r_k and mask of the tile cell PK, RK, NK
  W L[4], H[4];
  V PP[4], PH[4], PK[4], RK[4];
  uint8_t NK[4], M[2], HM[2];
  // q_{k-1}, r_k and p_k of plane q on the tile and the first ring (slot s,
  // x neighbours in slots sm (q - 1) and sp (q + 1), p_{k-1} plane pa[b]);
  // p_k to pk[bk]; the tile cell stored when q is in the run
  auto form = [&](int q, int s, int sm, int sp, int b, int bk) __attribute__((always_inline)) {
    {
      const W& w = L[s];
      V qp = V(0);
      if (!first)
        qp = stencil(w.nb, PP[s], PP[sp], pa[b][jl + 3][kl + 2], pa[b][jl + 2][kl + 3], PP[sm], pa[b][jl + 1][kl + 2],
                     pa[b][jl + 2][kl + 1]);
      const V rk = r_new(w, qp);
      const V pkv = p_new(w, rk);
      PK[s] = pkv;
      RK[s] = rk;
      NK[s] = w.nb;
      pk[bk][jl + 1][kl + 1] = pkv;
      if (q >= i_lo && q < i_hi) own(q, w, rk);
    }
    if (r1) {
      const W& w = H[s];
      V qp = V(0);
      if (!first)
        qp = stencil(w.nb, PH[s], PH[sp], pa[b][hj + 3][hk + 2], pa[b][hj + 2][hk + 3], PH[sm], pa[b][hj + 1][hk + 2],
                     pa[b][hj + 2][hk + 1]);
      pk[bk][hj + 1][hk + 1] = p_new(w, r_new(w, qp));
    }
  };
  // p_{k-1} of plane q (slot s) to pa[b]: tile cell and ring cell
  auto stage = [&](int s, int b) __attribute__((always_inline)) {
    PP[s] = pval(L[s]);
    pa[b][jl + 2][kl + 2] = PP[s];
    PH[s] = pval(H[s]);
    if (hr) pa[b][hj + 2][hk + 2] = PH[s];
  };
  auto load_plane = [&](int q, int s) __attribute__((always_inline)) {
    load_v(q, cok, load_m(q, cok, col), col, L[s], true, true);
    load_v(q, hok, load_m(q, hok, hcol), hcol, H[s], r1, false);
  };
  // ---- prologue: planes i_lo - 2 .. i_lo + 3 ----
  {
    W t2, t1, h1;
    load_v(i_lo - 2, cok, load_m(i_lo - 2, cok, col), col, t2, false, false);
    load_v(i_lo - 1, cok, load_m(i_lo - 1, cok, col), col, t1, true, false);
    load_v(i_lo - 1, hok && r1, load_m(i_lo - 1, hok && r1, hcol), hcol, h1, false, false);
    load_plane(i_lo, 0);
    load_plane(i_lo + 1, 1);
    // plane i_lo - 1: p_{k-1} of tile and first ring to pa[1]; its p_k (tile
    // cell only) is the -x neighbour of plane i_lo
    PP[3] = pval(t1);
    pa[1][jl + 2][kl + 2] = PP[3];
    PH[3] = pval(h1);
    if (r1) pa[1][hj + 2][hk + 2] = PH[3];
    PP[0] = pval(L[0]);
    lds_barrier();
    {
      V qp = V(0);
      if (!first)
        qp = stencil(t1.nb, PP[3], PP[0], pa[1][jl + 3][kl + 2], pa[1][jl + 2][kl + 3], pval(t2), pa[1][jl + 1][kl + 2],
                     pa[1][jl + 2][kl + 1]);
      PK[3] = p_new(t1, r_new(t1, qp));
    }
    load_plane(i_lo + 2, 2);
    load_plane(i_lo + 3, 3);
    M[0] = load_m(i_lo + 4, cok, col);
    HM[0] = load_m(i_lo + 4, hok, hcol);
    // plane i_lo: p_{k-1} to pa[0], then r_k, p_k (p_k to pk[0])
    stage(0, 0);
    lds_barrier();
    PP[1] = pval(L[1]);
    PH[1] = pval(H[1]);
    form(i_lo, 0, 3, 1, 0, 0);
    // plane i_lo + 1: p_{k-1} to pa[1]
    stage(1, 1);
    lds_barrier();
  }
  constexpr int UNR = 4;
  for (int ib = i_lo; ib < i_hi; ib += UNR) {
#pragma unroll
    for (int d = 0; d < UNR; ++d) {
      const int i = ib + d;
      if (i < i_hi) {  // uniform over the block
        const int s0 = d & 3, s1 = (d + 1) & 3, s2 = (d + 2) & 3, s3 = (d + 3) & 3;
        const int mb = d & 1;
        // masks of plane i + 5, vectors of plane i + 4 (into the slot of plane i)
        M[mb ^ 1] = load_m(i + 5, cok, col);
        HM[mb ^ 1] = load_m(i + 5, hok, hcol);
        load_v(i + 4, cok, M[mb], col, L[s0], true, true);
        load_v(i + 4, hok, HM[mb], hcol, H[s0], r1, false);
        // p_{k-1} of plane i + 2 to pa; r_k, p_k of plane i + 1
        stage(s2, d & 1);
        form(i + 1, s1, s0, s2, (d + 1) & 1, (d + 1) & 1);
        // q_k of plane i on the tile
        const uint8_t nb = NK[s0];
        if (nb & NB_UNKNOWN) {
          const int64_t pb = (int64_t)i * sx;
          const int bq = d & 1;
          const V pcv = PK[s0];
          const V qv = stencil(nb, pcv, PK[s1], pk[bq][jl + 2][kl + 1], pk[bq][jl + 1][kl + 2], PK[s3],
                               pk[bq][jl][kl + 1], pk[bq][jl + 1][kl]);
          const V rk = RK[s0];
          const V dv = d_of(nb);
          const V zv = rk * dv;
          st_nt(pout + pb + col, pcv);
          if (qout) st_nt(qout + pb + col, qv);
          a[2] = acc64(a[2], pcv, qv);
          a[3] = acc64(a[3], qv, zv);
          a[4] = acc64(a[4], qv, dv * qv);
          a[5] = acc64(a[5], qv, rk);
          a[6] = acc64(a[6], qv, qv);
        }
        lds_barrier();
      }
    }
  }
  write_partials<FDOT, SQ>(a, Pout);
}

// q_k = A p_k on the slow rows after k_cgf stored every p_k and r_k, with
// their share of the sums p.q, q.z, q.Dq, q.r, q.q (z = D r, D the stored
// slow-row diagonal); a fixed grid walks the list, so the sums are
// deterministic.
template <int OPk>
__global__ void __launch_bounds__(PB) k_slow_q5(OpArgs o, const uint32_t* __restrict__ list, int64_t n,
                                                const float* __restrict__ pc, const float* __restrict__ rk,
                                                float* __restrict__ q, double* __restrict__ S7, SolveScalars* sc,
                                                int k) {
  using V = float;
  if (sc->done) return;  // the solve stopped at launch k: no later launch reads these sums
  double a[FDOT] = {0, 0, 0, 0, 0, 0, 0};  // r.z, r.r are k_cgf's
  {
    const Grid& g = o.g;
    for (int64_t t = (int64_t)blockIdx.x * PB + threadIdx.x; t < n; t += (int64_t)gridDim.x * PB) {
      const int64_t idx = list[t];
      int i, j, kk;
      ijk_of(g, idx, i, j, kk);
      V pn[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int ii = i + c_off[f][0], jj = j + c_off[f][1], k2 = kk + c_off[f][2];
        pn[f] = valid(g, ii, jj, k2) ? pc[cid(g, ii, jj, k2)] : V(0);
      }
      const V pcv = pc[idx];
      const V qv = slow_row<V, OPk>(o, i, j, kk, idx, pcv, pn[0], pn[1], pn[2], pn[3], pn[4], pn[5]);
      q[idx] = qv;
      const V r = rk[idx];
      const V dv = o.precond ? dslow_at<V>(o, idx) : V(1);
      const V zv = r * dv;
      a[2] = acc64(a[2], pcv, qv);
      a[3] = acc64(a[3], qv, zv);
      a[4] = acc64(a[4], qv, dv * qv);
      a[5] = acc64(a[5], qv, r);
      a[6] = acc64(a[6], qv, qv);
    }
  }
  write_partials<FDOT>(a, S7);
}

// FP64 true residual of the slow rows (the tile pass skipped them), with the
// sums r.z, r.r, q.z, q.r written as blocks after the tile pass's partials.
template <int OPk>
__global__ void __launch_bounds__(PB) k_slow_resid64(OpArgs o, const uint32_t* __restrict__ list, int64_t n,
                                                     const float* __restrict__ b, const double* __restrict__ x64,
                                                     const float* __restrict__ x, float* __restrict__ r,
                                                     const float* __restrict__ qf, double* __restrict__ R4,
                                                     SolveScalars* sc) {
  if (sc->done) return;
  const Grid& g = o.g;
  double a[4] = {0, 0, 0, 0};
  for (int64_t t = (int64_t)blockIdx.x * PB + threadIdx.x; t < n; t += (int64_t)gridDim.x * PB) {
    const int64_t idx = list[t];
    int ci, cj, ck;
    ijk_of(g, idx, ci, cj, ck);
    const int8_t m = o.mark[idx];
    const double xc = x64[idx] + (double)x[idx];
    double sacc = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int ii = ci + c_off[f][0], jj = cj + c_off[f][1], kk = ck + c_off[f][2];
      const int8_t mn = mark_at(g, o.mark, ii, jj, kk);
      if (neumann(o, OPk, mn)) continue;
      const int64_t nbi = cid(g, ii, jj, kk);
      const double xn = is_unknown(OPk, mn) ? x64[nbi] + (double)x[nbi] : 0.0;
      const double bf = (OPk == OP_UNIT) ? 1.0 : face_beta<double>(o, idx, nbi, m, mn);
      sacc += bf * (xn - xc);
    }
    const double ax = sacc / ((double)g.h * (double)g.h);
    const double rv = (double)b[idx] - ax;
    const float rf = (float)rv;
    r[idx] = rf;
    const float dv = o.precond ? dslow_at<float>(o, idx) : 1.0f;
    a[0] = acc64(a[0], rf, rf * dv);
    a[1] += rv * rv;
    const float qv = qf[idx];
    a[2] = acc64(a[2], qv, rf * dv);
    a[3] = acc64(a[3], qv, rf);
  }
  write_partials<4>(a, R4);
}

// MIXED: x64 += x at exit.
// x += alpha_{k} p_{k} that update k left pending (alpha from its slot).
template <class V>
__global__ void __launch_bounds__(PB) k_flush_x(Launch L, int64_t G, const uint8_t* __restrict__ nbm,
                                                V* __restrict__ x, const V* __restrict__ p,
                                                const SolveScalars* __restrict__ sc, int slot) {
  const V al = (V)sc->al[slot];
  CELLS_OF_BLOCK(G) {
    if (!(nbm[idx] & NB_UNKNOWN)) continue;
    x[idx] = x[idx] + al * p[idx];
  }
}

__global__ void __launch_bounds__(PB) k_final_fold(Launch L, int64_t G, const uint8_t* __restrict__ nbm,
                                                   float* __restrict__ x, double* __restrict__ x64) {
  CELLS_OF_BLOCK(G) {
    if (!(nbm[idx] & NB_UNKNOWN)) continue;
    x64[idx] += (double)x[idx];
    x[idx] = 0.0f;
  }
}

// |b - A x|^2 in FP64 for the report (partials into S).
template <int OPk, class B, class X>
__global__ void __launch_bounds__(PB) k_true_resid(OpArgs o, Launch L, const B* __restrict__ b, const X* __restrict__ x,
                                                   Parts P) {
  double a_rr = 0;
  CELLS_OF_BLOCK(o.g.G) {
    const uint8_t nb = o.nbm[idx];
    if (!(nb & NB_UNKNOWN)) continue;
    double rv = (double)b[idx] - apply_op<double, OPk, X>(o, x, idx, nb);
    a_rr += rv * rv;
  }
  double v[1] = {a_rr};
  write_partials<1>(v, P.S);
}

// Sum of the S partials into sc->rr_true (one block).
__global__ void __launch_bounds__(PB) k_collect_rr(const double* __restrict__ S, int nb, SolveScalars* sc) {
  double t[1];
  read_partials<1>(S, nb, t);
  if (threadIdx.x == 0) sc->rr_true = t[0];
}

// r.r of a residual pass with four sums per block (the single-kernel form).
__global__ void __launch_bounds__(PB) k_collect_rr4(const double* __restrict__ S, int nb, SolveScalars* sc) {
  double t[4];
  read_partials<4>(S, nb, t);
  if (threadIdx.x == 0) sc->rr_true = t[1];
}

__global__ void k_clear_done(SolveScalars* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc->done = 0;
}

__global__ void k_to_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_to_double(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
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

// spmv tiles: TJ x TK (j, k) columns, planes split along i into runs so the
// launch has about `blocks` blocks: one resident round on 256 CUs, i.e. as
// many 512-thread blocks as fit a CU at the kernel's register use (three for
// FP32 vectors without slow rows, two otherwise).
Tiling make_tiling(const Grid& g, int blocks) {
  Tiling t;
  t.ntj = (g.ny + TJ - 1) / TJ;
  t.ntk = (g.nz + TK - 1) / TK;
  const int tiles = t.ntj * t.ntk;
  int nseg = blocks / tiles;
  if (nseg < 1) nseg = 1;
  if (nseg > g.nx) nseg = g.nx;
  t.seg = (g.nx + nseg - 1) / nseg;
  nseg = (g.nx + t.seg - 1) / t.seg;
  t.nblk = tiles * nseg;
  return t;
}

OpArgs make_op(Ctx& c, int precond, int opk) {
  OpArgs o;
  o.g = c.g;
  o.mark = c.mark;
  o.ls = c.lsmerge;
  o.nbm = c.nbm;
  o.dslow = c.dslow;
  o.il = 1.0 / (double)c.cfg.rho_liquid;
  o.ia = 1.0 / (double)c.cfg.rho_air;
  const double h2 = (double)c.g.h * (double)c.g.h;
  for (int t = 0; t < 6; ++t)
    o.hal[t] = c.cfg.heat_alpha[t] > 0 ? h2 / (double)c.cfg.heat_alpha[t] / (double)c.cfg.dt : 0.0;
  o.solid_c = c.cfg.solid_constraint;
  o.vac_neumann = c.cfg.heat_vacuum_neumann;
  o.precond = precond;
  // fast-path tables in each precision's own arithmetic
  const float hf = c.g.h, bef = (opk == OP_UNIT) ? 1.0f : (float)o.il;
  const double hd = (double)c.g.h, bed = (opk == OP_UNIT) ? 1.0 : o.il;
  o.dinv_f[0] = 1.0f;
  o.dinv_d[0] = 1.0;
  for (int n = 1; n <= 6; ++n) {
    o.dinv_f[n] = -(hf * hf) / ((float)n * bef);
    o.dinv_d[n] = -(hd * hd) / ((double)n * bed);
  }
  o.qc_f = bef / (hf * hf);
  o.qc_d = bed / (hd * hd);
  return o;
}

bool read_scalars(Ctx& c) {
  hipMemcpyAsync(c.scal_host, c.scal, sizeof(SolveScalars), hipMemcpyDeviceToHost, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "solver poll");
}

// Generic (P)CG driver.  V: vector type, Dacc: dot accumulation type.
// b: right-hand side (type V), x: solution (type V; MIXED adds x64).
// MIXED: after every refine_every iterations the FP64 true residual replaces
// the recurrence residual; convergence is accepted only when the residual
// that met the criterion was a true one.
template <class V, class Dacc, int OPk>
int cg_solve(Ctx& c, const V* b, V* x, double* x64, bool ref, int precond, int cap, double tol, bool mixed,
             mpf_solve_info* info) {
  const int64_t G = c.g.G;
  const Launch L = make_launch(G);
  const OpArgs o = make_op(c, precond, OPk);
  V* r = (V*)c.vr;
  V* pbuf[2] = {(V*)c.vp, (V*)c.vp2};
  V* q = (V*)c.vq;
  Parts P{c.partials, c.partials + 2 * MAX_BLK, c.partials + 4 * MAX_BLK, c.partials + 6 * MAX_BLK};
  const double eps_abs = 1e-2 * (double)G;
  const int R = c.cfg.refine_every > 0 ? c.cfg.refine_every : 20;
  int check = mixed ? R : c.cfg.check_every;
  if (check <= 0) check = 16;
  launch(c, "pcg_nbm", dim3(nblocks(G, 256)), dim3(256), 0, k_nbm<OPk, V>, o, c.nbm, (V*)c.dslow);
  // slow rows: list in cell order (positions from a scan into the cell-count scratch)
  launch(c, "pcg_slow_flag", dim3(nblocks(G, 256)), dim3(256), 0, k_slow_flag, G, (const uint8_t*)c.nbm, c.ctmp_u32);
  const int64_t nslow = run_scan_u32(c, c.ctmp_u32, c.count, G, true);
  if (nslow < 0) return MPF_HIP_ERROR;
  uint32_t* slow_list = c.ctmp_u32;
  if (nslow > 0)
    launch(c, "pcg_slow_list", dim3(nblocks(G, 256)), dim3(256), 0, k_slow_list, G, (const uint8_t*)c.nbm,
           (const uint32_t*)c.count, slow_list);
  // slow-row blocks: grid-stride over the list, at most MAX_BLK of them (the
  // single-kernel iteration sums their partials after its tile blocks, in a
  // region that holds 2 * MAX_BLK blocks)
  int nsb = (int)((nslow + PB - 1) / PB);
  if (nsb > MAX_BLK) nsb = MAX_BLK;
  auto spmv = nslow > 0 ? k_spmv<V, Dacc, OPk, true> : k_spmv<V, Dacc, OPk, false>;
  auto resid64 = nslow > 0 ? k_resid64<OPk, 1> : k_resid64<OPk, 0>;
  // the MIXED form without slow rows runs one kernel per iteration (k_cgf)
  const bool fused = std::is_same<V, float>::value && mixed && !ref;
  // without slow rows the fused iteration can form q_{k-1} again instead of
  // storing q (k_cgq, one 1024-thread block per CU; cfg.pcg_recompute_q)
  const bool qfree = fused && nslow == 0 && c.cfg.pcg_recompute_q;
  Tiling T = make_tiling(c.g, qfree ? 256 : fused ? 512 : ((sizeof(V) == 4 && nslow == 0) ? 768 : 512));
  T.ns5 = fused ? nsb : 0;
  // the true-residual pass keeps two 512-thread blocks per CU (its own
  // tiling when k_cgq runs one block per CU); its partial count is Tr.nblk
  Tiling Tr = qfree ? make_tiling(c.g, 512) : T;
  Tr.ns5 = T.ns5;
  // U-slot partials: update and init write L.nblk, the residual pass T.nblk;
  // both pad to npad, the count the spmv sums
  int npad = L.nblk > T.nblk ? L.nblk : T.nblk;
  if (Tr.nblk > npad) npad = Tr.nblk;
  if (npad > MAX_BLK) {
    c.err = "grid too large for the solver's reduction buffers";
    return MPF_INVALID_ARG;
  }
  launch(c, "pcg_init", dim3(L.nblk), dim3(PB), 0, k_init<V, Dacc, OPk>, o, L, b, x, r, pbuf[0], pbuf[1], (V*)c.vp3,
         mixed ? x64 : (double*)nullptr, P, c.scal, npad);
  int k = 0;
  bool fold_next = false;
  // MIXED restarts (the recurrence met the criterion, the FP64 true residual
  // did not): bounded, so a solve that cannot reach the tolerance in float
  // vectors stops with MPF_NO_CONVERGENCE instead of cycling
  int restarts = 0;
  bool true_missed = false;
  double prev_true_rr = -1.0;  // FP64 true r.r at the last restart (stagnation test)
  if constexpr (std::is_same<V, float>::value) {
    if (fused) {
      float* rb[2] = {(float*)c.vr, (float*)c.vr2};
      float* qb[2] = {(float*)c.vq, (float*)c.vq2};
      // p_k in pf[(k + 1) % 3]: a launch reads p_k-1 and p_k-2 and writes p_k
      float* pf[3] = {(float*)c.vp, (float*)c.vp2, (float*)c.vp3};
      // partial sums: per launch T.nblk tile blocks then T.ns5 slow-row blocks
      double* S7[2] = {c.partials + 8 * MAX_BLK, c.partials + 22 * MAX_BLK};
      double* Rs = c.partials + 36 * MAX_BLK;
      auto cg_iter = nslow > 0 ? k_cgf<OPk, true, false, 1> : k_cgf<OPk, false, false, 1>;
      auto cg_steady = nslow > 0 ? k_cgf<OPk, true, true, 1> : k_cgf<OPk, false, true, 1>;
      // without slow rows x advances every other launch (modes 2 and 3 of
      // k_cgf), and fully before each true-residual pass
      auto cg_skip = k_cgf<OPk, false, true, 2>;
      auto cg_double = k_cgf<OPk, false, true, 3>;
      auto cg_fold = k_cgf<OPk, false, false, 4>;  // x64 += x + alpha p after a true residual
      bool pending = false;  // the last launch left alpha p in x's debt
      auto res_tiles = nslow > 0 ? k_resid64<OPk, 2> : k_resid64<OPk, 0>;
      // FP64 true residual: the tile pass, then the slow rows' list pass
      auto true_resid = [&](int kk, float* rr_, const float* qq_) {
        launch(c, "pcg_resid64", dim3(Tr.nblk), dim3(ST), 0, res_tiles, o, Tr, (const float*)b, (const double*)x64,
               (const float*)x, rr_, P, c.scal, kk, npad, qq_, Rs);
        if (nslow > 0)
          launch(c, "pcg_slow_resid64", dim3(nsb), dim3(PB), 0, k_slow_resid64<OPk>, o, (const uint32_t*)slow_list,
                 nslow, (const float*)b, (const double*)x64, (const float*)x, rr_, qq_, Rs + 4 * Tr.nblk, c.scal);
      };
      int first = 1, fold = 0;
      bool use_res = false;
      while (k <= cap) {
        const int k_end = k + check;
        for (; k < k_end && k <= cap; ++k) {
          const double* pres = first == 1 ? (const double*)P.U1 : (first == 2 || use_res) ? Rs : nullptr;
          float* p_prev = pf[k % 3];
          float* p_next = pf[(k + 1) % 3];
          const float* p_prev2 = pf[(k + 2) % 3];
          if (qfree) {  // q_k stored only for the true-residual pass after this launch
            launch(c, "pcg_cg_iterq", dim3(T.nblk), dim3(SQ), 0, k_cgq<OPk>, o, T, Tr.nblk, first, npad,
                   (const double*)S7[(k + 1) & 1], pres, S7[k & 1], (const float*)rb[k & 1], (const float*)p_prev,
                   rb[(k + 1) & 1], p_next, (k + 1) % R == 0 ? qb[(k + 1) & 1] : (float*)nullptr, (float*)x, x64,
                   fold, c.scal, k, tol, cap);
          } else {
            auto kern = cg_iter;
            if (first == 0 && fold == 0) {
              kern = cg_steady;
              if (nslow == 0) {
                if (pending) {
                  kern = cg_double;
                  pending = false;
                } else if ((k + 1) % R != 0 && k < cap) {
                  kern = cg_skip;
                  pending = true;
                }
              }
            } else {
              pending = false;
              if (first == 0 && nslow == 0) kern = cg_fold;
            }
            launch(c, "pcg_cg_iter", dim3(T.nblk), dim3(ST), 0, kern, o, T, first, npad,
                   (const double*)S7[(k + 1) & 1], pres, S7[k & 1], (const float*)rb[k & 1],
                   (const float*)qb[k & 1], (const float*)p_prev, rb[(k + 1) & 1], qb[(k + 1) & 1], p_next,
                   (float*)x, x64, fold, c.scal, k, tol, cap, p_prev2);
          }
          if (nslow > 0)
            launch(c, "pcg_slow_q5", dim3(nsb), dim3(PB), 0, k_slow_q5<OPk>, o, (const uint32_t*)slow_list, nslow,
                   (const float*)p_next, (const float*)rb[(k + 1) & 1], qb[(k + 1) & 1], S7[k & 1] + FDOT * T.nblk,
                   c.scal, k);
          first = 0;
          fold = 0;
          use_res = false;
          if ((k + 1) % R == 0) {
            true_resid(k, rb[(k + 1) & 1], (const float*)qb[(k + 1) & 1]);
            use_res = true;
            fold = 1;
          }
        }
        if (!read_scalars(c)) return MPF_HIP_ERROR;
        const SolveScalars& s = *c.scal_host;
        if (!s.done) continue;
        if (!s.converged || s.nonfinite) break;
        // launch kd stopped and advanced x; convergence stands only on the
        // FP64 true residual of that x, else the iteration restarts from it
        const int kd = s.iter;
        launch(c, "pcg_clear", dim3(1), dim3(64), 0, k_clear_done, c.scal);
        true_resid(kd, rb[kd & 1], (const float*)qb[kd & 1]);
        launch(c, "pcg_collect", dim3(1), dim3(PB), 0, k_collect_rr4, (const double*)Rs, Tr.nblk + T.ns5, c.scal);
        if (!read_scalars(c)) return MPF_HIP_ERROR;
        const SolveScalars& t = *c.scal_host;
        if (t.rr_true == 0.0 || sqrt(t.rr_true) <= tol * sqrt(t.bb)) break;
        // stop at the cap, after too many restarts, or when a restart did
        // not at least halve the true residual (FP32 vectors at their limit)
        const bool stagnant = prev_true_rr >= 0.0 && t.rr_true > 0.25 * prev_true_rr;
        prev_true_rr = t.rr_true;
        if (kd >= cap || ++restarts > MAX_RESTARTS || stagnant) {
          true_missed = true;
          break;
        }
        first = 2;
        fold = 1;
        k = kd;
      }
    }
  }
  if (!fused) {
    // x advances every other iteration (update modes: see k_update); half[k]
    // records that update k left alpha_k p_k pending in x
    std::vector<char> half((size_t)cap + 2, 0);
    auto flush = [&](int kd) {  // x += alpha_{kd-1} p_{kd-1} if update kd-1 left it pending
      if (kd < 1 || !half[kd - 1]) return;
      launch(c, "pcg_flush_x", dim3(L.nblk), dim3(PB), 0, k_flush_x<V>, L, G, (const uint8_t*)c.nbm, x,
             (const V*)pbuf[kd & 1], (const SolveScalars*)c.scal, (kd - 1) & 1);
      half[kd - 1] = 0;
    };
    while (k <= cap) {
      const int k_end = k + check;
      for (; k < k_end && k <= cap; ++k) {
        launch(c, "pcg_spmv", dim3(T.nblk), dim3(ST), 0, spmv, o, T, npad, (const V*)r, (const V*)pbuf[k & 1],
               pbuf[(k + 1) & 1], q, P, c.scal, k, (int)ref, eps_abs, tol, cap);
        if (nsb > 0)
          launch(c, "pcg_slow_q", dim3(nsb), dim3(PB), 0, k_slow_q<V, Dacc, OPk>, o, (const uint32_t*)slow_list, nslow,
                 (const V*)pbuf[(k + 1) & 1], q, P.S2, c.scal);
        const bool resid_next = mixed && (k + 1) % R == 0;
        const bool h = (k % 2 == 0) && !resid_next && !fold_next;
        const int mode = h ? 0 : ((k > 0 && half[k - 1]) ? 2 : 1);
        half[k] = h;
        launch(c, "pcg_update", dim3(L.nblk), dim3(PB), 0, k_update<V, Dacc, OPk>, o, L, x, r,
               (const V*)pbuf[(k + 1) & 1], (const V*)pbuf[k & 1], (const V*)q, mixed ? x64 : (double*)nullptr, P,
               T.nblk, nsb, c.scal, k, (int)fold_next, mode, npad);
        fold_next = false;
        if (mixed && (k + 1) % R == 0) {
          launch(c, "pcg_resid64", dim3(T.nblk), dim3(ST), 0, resid64, o, T, (const float*)b,
                 (const double*)x64, (const float*)x, (float*)r, P, c.scal, k, npad, (const float*)nullptr,
                 (double*)nullptr);
          fold_next = true;
        }
      }
      if (!read_scalars(c)) return MPF_HIP_ERROR;
      const SolveScalars& s = *c.scal_host;
      if (!s.done) continue;
      // updates after s.iter - 1 returned without running: x takes what the
      // last one that ran left pending
      flush(s.iter);
      if (!mixed || !s.converged || s.nonfinite) break;
      // spmv(kd) met the criterion on the residual of update(kd-1).  Which
      // true-residual pass last ran is read from the device: passes launched
      // after the solve finished returned without running.
      const int kd = s.iter;
      if (kd == 0 || s.last_true == kd - 1) break;
      if (kd >= cap || ++restarts > MAX_RESTARTS) {
        true_missed = true;
        break;
      }
      // resid64 returns early while done is set: clear it first
      launch(c, "pcg_clear", dim3(1), dim3(64), 0, k_clear_done, c.scal);
      launch(c, "pcg_resid64", dim3(T.nblk), dim3(ST), 0, resid64, o, T, (const float*)b, (const double*)x64,
             (const float*)x, (float*)r, P, c.scal, kd - 1, npad, (const float*)nullptr, (double*)nullptr);
      fold_next = true;
      k = kd;
    }
  }
  if (mixed) {
    launch(c, "pcg_fold", dim3(L.nblk), dim3(PB), 0, k_final_fold, L, G, (const uint8_t*)c.nbm, (float*)x, x64);
    launch(c, "pcg_true_resid", dim3(L.nblk), dim3(PB), 0, k_true_resid<OPk, V, double>, o, L, b,
           (const double*)x64, P);
  } else {
    launch(c, "pcg_true_resid", dim3(L.nblk), dim3(PB), 0, k_true_resid<OPk, V, V>, o, L, b, (const V*)x, P);
  }
  launch(c, "pcg_collect", dim3(1), dim3(PB), 0, k_collect_rr, (const double*)P.S, L.nblk, c.scal);
  if (!read_scalars(c)) return MPF_HIP_ERROR;
  const SolveScalars& s = *c.scal_host;
  const bool converged = s.converged && !true_missed;
  if (info) {
    info->iterations = s.iter;
    info->refinements = s.refinements;
    info->converged = converged ? 1 : 0;
    info->rel_residual = s.bb > 0 ? sqrt(s.rr_true) / sqrt(s.bb) : 0.0;
    info->abs_rr = s.rr;
  }
  if (s.nonfinite) {
    c.err = "non-finite value in the solver";
    return MPF_NONFINITE;
  }
  if (true_missed)
    c.note = "mixed-precision solve: the FP64 true residual missed the tolerance (" + std::to_string(restarts) +
             " restarts, iteration " + std::to_string(s.iter) + ")";
  return converged ? MPF_OK : MPF_NO_CONVERGENCE;
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

// A x = rhs for the adjoint of the projection (the operator is symmetric), in
// the configured precision; rhs and the solution are FP64 cell fields.
int run_adjoint_solve(Ctx& c, const double* rhs, double* sol, mpf_solve_info* info) {
  const mpf_config& k = c.cfg;
  const bool ref = k.solver == MPF_SOLVER_REFERENCE_CG;
  const int precond = ref ? 0 : 1;
  const int cap = ref ? k.ref_max_iters : k.max_iters;
  const bool vd = k.op == MPF_OPERATOR_VARDENS;
  if (k.precision == MPF_FP64) {
    if (vd) return cg_solve<double, double, OP_VARDENS>(c, rhs, sol, nullptr, ref, precond, cap, k.tol, false, info);
    return cg_solve<double, double, OP_UNIT>(c, rhs, sol, nullptr, ref, precond, cap, k.tol, false, info);
  }
  // FP32 vectors: the right-hand side rounded to float into the scratch field
  const int64_t G = c.g.G;
  launch(c, "adj_to_float", dim3(nblocks(G, 256)), dim3(256), 0, k_to_float, G, rhs, c.ctmp);
  if (k.precision == MPF_MIXED) {
    if (vd) return cg_solve<float, double, OP_VARDENS>(c, (const float*)c.ctmp, (float*)c.vx, sol, ref, precond, cap, k.tol, true, info);
    return cg_solve<float, double, OP_UNIT>(c, (const float*)c.ctmp, (float*)c.vx, sol, ref, precond, cap, k.tol, true, info);
  }
  int st = vd ? cg_solve<float, float, OP_VARDENS>(c, (const float*)c.ctmp, (float*)c.vx, nullptr, ref, precond, cap, k.tol, false, info)
              : cg_solve<float, float, OP_UNIT>(c, (const float*)c.ctmp, (float*)c.vx, nullptr, ref, precond, cap, k.tol, false, info);
  launch(c, "adj_to_double", dim3(nblocks(G, 256)), dim3(256), 0, k_to_double, G, (const float*)c.vx, sol);
  return st;
}

// Heat: reference CG (no preconditioner, r.r < 1e-2 G), FP32 vectors, FP64 dots.
int run_heat_cg(Ctx& c, mpf_solve_info* info) {
  return cg_solve<float, double, OP_HEAT>(c, (const float*)c.ctmp, (float*)c.tp, nullptr, true, 0, c.cfg.heat_max_iters,
                                          0.0, false, info);
}

}  // namespace mpf
