// Cell hash, stable counting sort and reorder (a1-a3).
//
// Replaces cspray::hashAndSortParticles (spray.cu:614-637): calcHashD
// (spray_k.cu:1010-1028) + thrust::sort_by_key (spray_thrust.cu:12-17) +
// reorderDataAndFindCellStartD (spray_k.cu:1052-1129).  Result contract (the
// reference's, thrust's radix sort being stable for uint keys): particles
// ordered by (cell key, original index), gridHash = sorted keys, gridIndex =
// the permutation, cell_start/cell_end per cell (MPF_CELL_UNDEF when empty),
// all eight attribute arrays permuted.
//
// Counting sort instead of a comparison/radix sort, since keys are cell ids:
//  1. hash + count: one atomic per run of equal keys in a wavefront (input is
//     nearly sorted from the previous step, so runs are long);
//  2. exclusive scan of the per-cell counts (reduce-then-scan, LDS tiles);
//  3. scatter: one atomic per run reserves slots, lanes of a run take
//     consecutive slots in index order;
//  4. per-cell segment sort of the permutation (runs from different waves can
//     interleave within a cell): a register sorting network for <=16 entries;
//  5. gather of the 45-byte particle record into the other buffer.
#include <mutex>

#include "mpf_internal.h"

namespace mpf {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// getijkfrompos (spray_k.cu:51-57) + getidx (spray_k.cu:66-69): IEEE division,
// per-component range test, truncation.
__device__ __forceinline__ uint32_t hash_key(const Grid& g, f3 p) {
  float qx = (p.x - g.gx0) / g.h, qy = (p.y - g.gy0) / g.h, qz = (p.z - g.gz0) / g.h;
  int i = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
  int j = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
  int k = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
  return (uint32_t)cid(g, i, j, k);
}

// Run-length of equal keys starting at a head lane.  Active lanes form a
// prefix of the wavefront (particle indices are contiguous).
__device__ __forceinline__ void wave_runs(bool act, uint32_t k, bool& head, int& runlen, int& headlane) {
  const int lane = lane_id();
  uint32_t prev = __shfl_up(k, 1);
  head = act && (lane == 0 || prev != k);
  unsigned long long heads = __ballot(head);
  unsigned long long actm = __ballot(act);
  int nact = __popcll(actm);
  unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  unsigned long long above = heads & ~le;
  int next = above ? (__ffsll((long long)above) - 1) : 64;
  runlen = min(next, nact) - lane;
  unsigned long long below = heads & le;
  headlane = below ? (63 - __clzll((long long)below)) : 0;
}

__global__ void k_hash_count(Grid g, const f3* __restrict__ pos, int64_t P, uint32_t* __restrict__ key,
                             uint32_t* __restrict__ count) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = p < P;
  uint32_t k = 0xffffffffu;
  if (act) {
    k = hash_key(g, pos[p]);
    key[p] = k;
  }
  bool head;
  int runlen, headlane;
  wave_runs(act, k, head, runlen, headlane);
  if (head) atomicAdd(&count[k], (uint32_t)runlen);
}

__global__ void k_scatter(int64_t P, const uint32_t* __restrict__ key, uint32_t* __restrict__ cursor,
                          uint32_t* __restrict__ perm, uint32_t* __restrict__ skey) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = p < P;
  uint32_t k = act ? key[p] : 0xffffffffu;
  bool head;
  int runlen, headlane;
  wave_runs(act, k, head, runlen, headlane);
  uint32_t base = 0;
  if (head) base = atomicAdd(&cursor[k], (uint32_t)runlen);
  base = __shfl(base, headlane);
  if (act) {
    uint32_t slot = base + (uint32_t)(lane_id() - headlane);
    perm[slot] = (uint32_t)p;
    skey[slot] = k;
  }
}

// Keyless variants for mpf_hash_sort_noalloc: the count goes straight into
// the caller's cell_end, and the scatter recomputes each key from the
// position (no per-particle key array).
__global__ void k_hash_count_nokey(Grid g, const f3* __restrict__ pos, int64_t P, uint32_t* __restrict__ count) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = p < P;
  uint32_t k = act ? hash_key(g, pos[p]) : 0xffffffffu;
  bool head;
  int runlen, headlane;
  wave_runs(act, k, head, runlen, headlane);
  if (head) atomicAdd(&count[k], (uint32_t)runlen);
}

__global__ void k_scatter_rehash(Grid g, const f3* __restrict__ pos, int64_t P, uint32_t* __restrict__ cursor,
                                 uint32_t* __restrict__ perm, uint32_t* __restrict__ skey) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = p < P;
  uint32_t k = act ? hash_key(g, pos[p]) : 0xffffffffu;
  bool head;
  int runlen, headlane;
  wave_runs(act, k, head, runlen, headlane);
  uint32_t base = 0;
  if (head) base = atomicAdd(&cursor[k], (uint32_t)runlen);
  base = __shfl(base, headlane);
  if (act) {
    uint32_t slot = base + (uint32_t)(lane_id() - headlane);
    perm[slot] = (uint32_t)p;
    skey[slot] = k;
  }
}

__global__ void k_fill_u32(int64_t n, uint32_t* __restrict__ v, uint32_t x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = x;
}

// counts (in cend) + exclusive scan (in cstart) -> cstart = start or UNDEF,
// cend = start as the scatter cursor (UNDEF when empty); the scatter then
// advances each non-empty cell's cursor to its end.
__global__ void k_cells_cursor(int64_t G, uint32_t* __restrict__ cstart, uint32_t* __restrict__ cend) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const uint32_t n = cend[c], s = cstart[c];
  cstart[c] = n ? s : MPF_CELL_UNDEF;
  cend[c] = n ? s : MPF_CELL_UNDEF;
}

template <int N>
__device__ __forceinline__ void bitonic(uint32_t (&a)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        int l = i ^ j;
        if (l > i) {
          uint32_t x = a[i], y = a[l];
          bool up = (i & k) == 0;
          bool sw = up ? (x > y) : (x < y);
          a[i] = sw ? y : x;
          a[l] = sw ? x : y;
        }
      }
    }
  }
}

// A slice of n <= N entries: read together (past the end: the last entry),
// checked for order (runs usually arrive in order) and sorted in registers
// when not.
template <int N>
__device__ __forceinline__ void seg_small(uint32_t* __restrict__ v, uint32_t n) {
  uint32_t a[N];
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = v[(uint32_t)q < n ? q : n - 1];
  bool ordered = true;
#pragma unroll
  for (int q = 1; q < N; ++q) ordered = ordered && ((uint32_t)q >= n || a[q - 1] < a[q]);
  if (ordered) return;
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = ((uint32_t)q < n) ? a[q] : 0xffffffffu;
  bitonic<N>(a);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if ((uint32_t)q < n) v[q] = a[q];
}

// Sort each cell's slice of the permutation by particle index.
__global__ void k_segsort(int64_t G, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                          uint32_t* __restrict__ perm) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  uint32_t s = cstart[c];
  if (s == MPF_CELL_UNDEF) return;
  uint32_t n = cend[c] - s;
  if (n < 2) return;
  uint32_t* v = perm + s;
  if (n <= 8) seg_small<8>(v, n);
  else if (n <= 16) seg_small<16>(v, n);
  else if (n <= 32) seg_small<32>(v, n);
  else {
    for (uint32_t q = 1; q < n; ++q) {
      uint32_t x = v[q];
      uint32_t r = q;
      while (r > 0 && v[r - 1] > x) { v[r] = v[r - 1]; --r; }
      v[r] = x;
    }
  }
}

__global__ void k_reorder(int64_t P, const uint32_t* __restrict__ perm,
                          const f3* __restrict__ ipos, const f3* __restrict__ ivel, const int8_t* __restrict__ iflag,
                          const float* __restrict__ imass, const float* __restrict__ itemp,
                          const float* __restrict__ ilheat, const float* __restrict__ isolub,
                          const float* __restrict__ igas,
                          f3* __restrict__ opos, f3* __restrict__ ovel, int8_t* __restrict__ oflag,
                          float* __restrict__ omass, float* __restrict__ otemp, float* __restrict__ olheat,
                          float* __restrict__ osolub, float* __restrict__ ogas) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P) return;
  uint32_t p = perm[s];
  opos[s] = ipos[p];
  ovel[s] = ivel[p];
  oflag[s] = iflag[p];
  omass[s] = imass[p];
  otemp[s] = itemp[p];
  olheat[s] = ilheat[p];
  osolub[s] = isolub[p];
  ogas[s] = igas[p];
}

// counts -> cell_start / cell_end (UNDEF when empty) and scatter cursors.
__global__ void k_cells_from_scan(int64_t G, uint32_t* __restrict__ count, uint32_t* __restrict__ cstart,
                                  uint32_t* __restrict__ cend) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  uint32_t n = count[c], s = cstart[c];
  count[c] = s;
  cstart[c] = n ? s : MPF_CELL_UNDEF;
  cend[c] = n ? s + n : MPF_CELL_UNDEF;
}

__global__ void k_cell_mask(int64_t G, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                            const int8_t* __restrict__ flag, uint8_t* __restrict__ cmask) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const uint32_t s = cstart[c], e = cend[c];
  uint8_t m = 0;
  if (s != MPF_CELL_UNDEF) {
    m = M_ANY;
    // eight flags read together (past the end: the first again; the mask is
    // an OR, so repeats change nothing)
    for (uint32_t p = s; p < e; p += 8) {
      int8_t f8[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f8[u] = flag[p + u < e ? p + u : s];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int8_t f = f8[u];
        if (f == FLUID || f == SOLID) m |= M_FLUIDSOLID;
        if (f == SOLID) m |= M_SOLID;
        if (f == AIR) m |= M_AIR;
        if (f == AIRSOLO) m |= M_AIRSOLO;
      }
    }
  }
  cmask[c] = m;
}

// Cells whose list does not index the particle array: start != undefined and
// not start < end <= P.  Caller-supplied lists are checked before any gather
// walks them.
__global__ void k_cell_check(int64_t G, int64_t P, const uint32_t* __restrict__ cstart,
                             const uint32_t* __restrict__ cend, uint32_t* __restrict__ bad) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const uint32_t s = cstart[c], e = cend[c];
  const bool ok = (s == MPF_CELL_UNDEF) || (s < e && (int64_t)e <= P);
  if (!ok) atomicAdd(bad, 1u);
}

// ---- exclusive scan of uint32 (reduce, scan of block sums, apply) ----------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across the block (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int q = 0; q < SCAN_THREADS / 64; ++q) {
    if (q < w) off += wsum[q];
    tot += wsum[q];
  }
  total = tot;
  __syncthreads();
  return off + inc - v;
}

__global__ void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ blk) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int q = 0; q < SCAN_PER_THREAD; ++q) {
    int64_t e = base + (int64_t)q * SCAN_THREADS + threadIdx.x;
    if (e < n) s += in[e];
  }
  uint32_t total;
  block_excl_scan(s, wsum, total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// One block scans nb block sums in place (exclusive) and writes the total at blk[nb].
__global__ void k_scan_blocks(uint32_t* blk, int64_t nb) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_THREADS) {
    int64_t e = b0 + threadIdx.x;
    uint32_t v = (e < nb) ? blk[e] : 0;
    uint32_t total;
    uint32_t ex = block_excl_scan(v, wsum, total);
    if (e < nb) blk[e] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) blk[nb] = carry;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ in, int64_t n, const uint32_t* __restrict__ blk,
                             uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[SCAN_TILE + SCAN_TILE / 16];
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int q = 0; q < SCAN_PER_THREAD; ++q) {
    int e = q * SCAN_THREADS + threadIdx.x;
    int64_t ge = base + e;
    tile[e + (e >> 4)] = (ge < n) ? in[ge] : 0u;
  }
  __syncthreads();
  uint32_t loc[SCAN_PER_THREAD];
  uint32_t run = 0;
  const int t0 = threadIdx.x * SCAN_PER_THREAD;
#pragma unroll
  for (int q = 0; q < SCAN_PER_THREAD; ++q) {
    int e = t0 + q;
    loc[q] = run;
    run += tile[e + (e >> 4)];
  }
  uint32_t total;
  uint32_t off = block_excl_scan(run, wsum, total) + blk[blockIdx.x];
#pragma unroll
  for (int q = 0; q < SCAN_PER_THREAD; ++q) {
    int e = t0 + q;
    tile[e + (e >> 4)] = loc[q] + off;
  }
  __syncthreads();
  for (int q = 0; q < SCAN_PER_THREAD; ++q) {
    int e = q * SCAN_THREADS + threadIdx.x;
    int64_t ge = base + e;
    if (ge < n) out[ge] = tile[e + (e >> 4)];
  }
}

}  // namespace

__global__ void k_iota(int64_t n, uint32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

int64_t run_scan_u32(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n, bool want_total) {
  if (n <= 0) return 0;
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb + 1 > c.scan_blk_n) {
    c.err = "scan scratch too small";
    return -1;
  }
  launch(c, "scan_reduce", dim3((unsigned)nb), dim3(SCAN_THREADS), 0, k_scan_reduce, in, n, c.scan_blk);
  launch(c, "scan_blocks", dim3(1), dim3(SCAN_THREADS), 0, k_scan_blocks, c.scan_blk, nb);
  launch(c, "scan_apply", dim3((unsigned)nb), dim3(SCAN_THREADS), 0, k_scan_apply, in, n,
         (const uint32_t*)c.scan_blk, out);
  if (!want_total) return 0;
  uint32_t tot = 0;
  hipMemcpyAsync(c.counters_host + 15, c.scan_blk + nb, 4, hipMemcpyDeviceToHost, c.stream);
  hipStreamSynchronize(c.stream);
  tot = c.counters_host[15];
  return (int64_t)tot;
}

int64_t run_check_cells(Ctx& c) {
  hipMemsetAsync(c.counters + 6, 0, 4, c.stream);
  launch(c, "cell_check", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_cell_check, c.g.G, c.P,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, c.counters + 6);
  hipMemcpyAsync(c.counters_host + 6, c.counters + 6, 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "cell list check")) return -1;
  return c.counters_host[6];
}

void run_cell_mask(Ctx& c) {
  const int64_t G = c.g.G;
  launch(c, "cell_mask", dim3(nblocks(G, 256)), dim3(256), 0, k_cell_mask, G, (const uint32_t*)c.cstart,
         (const uint32_t*)c.cend, (const int8_t*)c.flag[c.cur], c.cmask);
  run_bricks(c);
}

void run_hash_sort(Ctx& c) {
  const int64_t P = c.P, G = c.g.G;
  const int cur = c.cur, nxt = 1 - c.cur;
  if (c.sorted && c.own_sort) {
    // nothing moved since this library's last sort: the stable sort is the
    // identity (keys, cell ranges and order stand); flags may have changed,
    // so the occupancy masks are rebuilt
    launch(c, "sort_identity", dim3(nblocks(P, 256)), dim3(256), 0, k_iota, P, c.perm);
    run_cell_mask(c);
    return;
  }
  hipMemsetAsync(c.count, 0, (size_t)G * 4, c.stream);
  launch(c, "hash_count", dim3(nblocks(P, 256)), dim3(256), 0, k_hash_count, c.g, (const f3*)c.pos[cur], P,
         c.key, c.count);
  run_scan_u32(c, c.count, c.cstart, G, false);
  launch(c, "cells_from_scan", dim3(nblocks(G, 256)), dim3(256), 0, k_cells_from_scan, G, c.count, c.cstart,
         c.cend);
  launch(c, "scatter", dim3(nblocks(P, 256)), dim3(256), 0, k_scatter, P, (const uint32_t*)c.key, c.count,
         c.perm, c.ghash);
  launch(c, "segsort", dim3(nblocks(G, 256)), dim3(256), 0, k_segsort, G, (const uint32_t*)c.cstart,
         (const uint32_t*)c.cend, c.perm);
  launch(c, "reorder", dim3(nblocks(P, 256)), dim3(256), 0, k_reorder, P, (const uint32_t*)c.perm,
         (const f3*)c.pos[cur], (const f3*)c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.mass[cur],
         (const float*)c.temp[cur], (const float*)c.lheat[cur], (const float*)c.solub[cur],
         (const float*)c.gas[cur], c.pos[nxt], c.vel[nxt], c.flag[nxt], c.mass[nxt], c.temp[nxt], c.lheat[nxt],
         c.solub[nxt], c.gas[nxt]);
  c.cur = nxt;
  c.user_cstart = c.user_cend = false;
  c.sorted = true;
  c.own_sort = true;
  c.mc_sorted = false;
  run_cell_mask(c);
}

// cspray::preMC / hashAndSortParticles_MC (spray.cu:639-662, 759-765): the
// same hash, stable sort and reorder on the 2N grid (calcHashD_MC,
// spray_k.cu:1030-1048, cell size (dx / 2nx) * nx); its cell lists go to the
// mc_* arrays, and the N-grid lists are stale afterwards.
void run_hash_sort_mc(Ctx& c) {
  const int64_t P = c.P, G2 = c.g2.G;
  const int cur = c.cur, nxt = 1 - c.cur;
  hipMemsetAsync(c.mc_count, 0, (size_t)G2 * 4, c.stream);
  launch(c, "mc_hash_count", dim3(nblocks(P, 256)), dim3(256), 0, k_hash_count, c.g2, (const f3*)c.pos[cur], P,
         c.key, c.mc_count);
  run_scan_u32(c, c.mc_count, c.mc_cstart, G2, false);
  launch(c, "mc_cells_from_scan", dim3(nblocks(G2, 256)), dim3(256), 0, k_cells_from_scan, G2, c.mc_count,
         c.mc_cstart, c.mc_cend);
  launch(c, "mc_scatter", dim3(nblocks(P, 256)), dim3(256), 0, k_scatter, P, (const uint32_t*)c.key, c.mc_count,
         c.perm, c.ghash);
  launch(c, "mc_segsort", dim3(nblocks(G2, 256)), dim3(256), 0, k_segsort, G2, (const uint32_t*)c.mc_cstart,
         (const uint32_t*)c.mc_cend, c.perm);
  launch(c, "mc_reorder", dim3(nblocks(P, 256)), dim3(256), 0, k_reorder, P, (const uint32_t*)c.perm,
         (const f3*)c.pos[cur], (const f3*)c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.mass[cur],
         (const float*)c.temp[cur], (const float*)c.lheat[cur], (const float*)c.solub[cur],
         (const float*)c.gas[cur], c.pos[nxt], c.vel[nxt], c.flag[nxt], c.mass[nxt], c.temp[nxt], c.lheat[nxt],
         c.solub[nxt], c.gas[nxt]);
  c.cur = nxt;
  c.user_cstart = c.user_cend = false;
  c.sorted = false;
  c.own_sort = false;
  c.mc_sorted = true;
}

// ---- mpf_hash_sort_noalloc: the sort without a context -------------------
// Device scratch of the scan's block sums: a file-scope array, one copy per
// device, held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in); it looks up the
// device address once per device and passes that.
constexpr int64_t NOALLOC_SCAN_BLOCKS = 65536;
constexpr int64_t NOALLOC_MAX_CELLS = (NOALLOC_SCAN_BLOCKS - 1) * SCAN_TILE;
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCAN_BLOCKS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
uint32_t* noalloc_scratch(int dev, std::string& err) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    err = "device index out of range for the no-allocation scratch";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = (uint32_t*)ptr;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace mpf

using namespace mpf;

extern "C" mpf_status mpf_hash_sort_noalloc(const mpf_config* cfg, int64_t n, const float* pos, uint32_t* grid_hash,
                                            uint32_t* grid_index, uint32_t* cell_start, uint32_t* cell_end,
                                            void* stream) {
  auto fail = [](mpf_status st, const std::string& m) {
    g_noctx_err = m;
    return st;
  };
  if (!cfg || !cell_start || !cell_end) return fail(MPF_INVALID_ARG, "null argument");
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || !(cfg->dx > 0))
    return fail(MPF_INVALID_ARG, "invalid grid");
  const int64_t G = (int64_t)cfg->nx * cfg->ny * cfg->nz;
  if (G > NOALLOC_MAX_CELLS)
    return fail(MPF_INVALID_ARG, "grid of " + std::to_string(G) + " cells exceeds the no-allocation limit of " +
                                     std::to_string(NOALLOC_MAX_CELLS));
  if (n < 0 || n > 0xfffffff0ll) return fail(MPF_INVALID_ARG, "particle count out of range");
  if (n > 0 && (!pos || !grid_hash || !grid_index)) return fail(MPF_INVALID_ARG, "null particle array");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(MPF_HIP_ERROR, "hipGetDevice failed");
  std::string err;
  uint32_t* blk = noalloc_scratch(dev, err);
  if (!blk) return fail(MPF_HIP_ERROR, err);
  Grid g{};
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.G = G;
  g.h = cfg->dx;
  g.gx0 = cfg->gmin[0]; g.gy0 = cfg->gmin[1]; g.gz0 = cfg->gmin[2];
  g.dnz = make_fastdiv((uint32_t)g.nz);
  g.dny = make_fastdiv((uint32_t)g.ny);
  hipStream_t st = (hipStream_t)stream;
  const f3* p3 = (const f3*)pos;
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(G, 256)), dim3(256), 0, st, G, cell_start, MPF_CELL_UNDEF);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(G, 256)), dim3(256), 0, st, G, cell_end, MPF_CELL_UNDEF);
  } else {
    const int64_t nb = (G + SCAN_TILE - 1) / SCAN_TILE;
    // counts in cell_end (cleared by a kernel, not a memset: graph-capture safe)
    hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(G, 256)), dim3(256), 0, st, G, cell_end, 0u);
    hipLaunchKernelGGL(k_hash_count_nokey, dim3(nblocks(n, 256)), dim3(256), 0, st, g, p3, n, cell_end);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, (const uint32_t*)cell_end, G,
                       blk);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_THREADS), 0, st, blk, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, (const uint32_t*)cell_end, G,
                       (const uint32_t*)blk, cell_start);
    hipLaunchKernelGGL(k_cells_cursor, dim3(nblocks(G, 256)), dim3(256), 0, st, G, cell_start, cell_end);
    hipLaunchKernelGGL(k_scatter_rehash, dim3(nblocks(n, 256)), dim3(256), 0, st, g, p3, n, cell_end, grid_index,
                       grid_hash);
    hipLaunchKernelGGL(k_segsort, dim3(nblocks(G, 256)), dim3(256), 0, st, G, (const uint32_t*)cell_start,
                       (const uint32_t*)cell_end, grid_index);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MPF_HIP_ERROR, std::string("kernel launch: ") + hipGetErrorString(e));
  return MPF_OK;
}
