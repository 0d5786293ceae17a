// Internal declarations of the MI355X FLIP multiphase library (gfx950).
//
// Device code follows the per-element expressions of the reference kernels
// (Multiphase/Multiphase/spray_k.cu) in the same operation order; the library
// is built with -ffp-contract=off so those expressions round exactly like the
// CPU oracle wherever an integer (hash key, flag, case index) is derived from
// floating point.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mpflip.h"

namespace mpf {

constexpr int8_t FLUID = MPF_TYPEFLUID, AIR = MPF_TYPEAIR, BOUNDARY = MPF_TYPEBOUNDARY,
                 VACUUM = MPF_TYPEVACUUM, AIRSOLO = MPF_TYPEAIRSOLO, SOLID = MPF_TYPESOLID;

// Per-cell occupancy mask built after every sort (bit set when the cell holds
// at least one particle of that kind); lets gathers skip cells cheaply.
constexpr uint8_t M_FLUIDSOLID = 1, M_AIR = 2, M_AIRSOLO = 4, M_SOLID = 8, M_ANY = 16;

struct f3 {
  float x, y, z;
};

// Division of a 32-bit unsigned numerator by a run-time-invariant divisor
// with one multiply-high, an add and a shift (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, the
// round-up variant): q = (mulhi(n, m) + n) >> s, exact for every n < 2^32.
// Replaces the 64-bit divisions of index decomposition in every kernel.
struct FastDiv {
  uint32_t d, m, s;
};
__host__ inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  f.s = s;
  // m = floor(2^32 * (2^s - d) / d) + 1, fits in 32 bits for d >= 1
  f.m = (uint32_t)(((((uint64_t)1 << s) - d) << 32) / d + 1);
  return f;
}
__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t = __umulhi(n, f.m);
#else
  uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
  return (uint32_t)(((uint64_t)t + n) >> f.s);
}

// Everything a kernel needs about the grid and the configuration; passed by
// value as a kernel argument (replaces __constant__ FlipConstant dparam and
// NX/NY/NZ, spray_k.cu:8-14).
struct Grid {
  int nx, ny, nz;
  int64_t G, FX, FY, FZ;
  float h;
  float gx0, gy0, gz0;  // gmin
  float gx1, gy1, gz1;  // gmax
  FastDiv dnz, dny, dnz1, dny1;  // nz, ny, nz+1, ny+1
};

__host__ __device__ inline int64_t cid(const Grid& g, int i, int j, int k) {
  return ((int64_t)i * g.ny + j) * g.nz + k;
}
__host__ __device__ inline bool valid(const Grid& g, int i, int j, int k) {
  return !(i < 0 || i > g.nx - 1 || j < 0 || j > g.ny - 1 || k < 0 || k > g.nz - 1);
}
// idx -> (i, j, k) of a cell (all grids here have fewer than 2^32 cells/faces).
__host__ __device__ inline void ijk_of(const Grid& g, int64_t idx, int& i, int& j, int& k) {
  uint32_t n = (uint32_t)idx;
  uint32_t r = fdiv(n, g.dnz);
  k = (int)(n - r * (uint32_t)g.nz);
  uint32_t q = fdiv(r, g.dny);
  j = (int)(r - q * (uint32_t)g.ny);
  i = (int)q;
}
// idx -> (i, j, k) of face axis A (dims (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1)).
template <int A>
__host__ __device__ inline void ijk_face(const Grid& g, int64_t idx, int& i, int& j, int& k) {
  const FastDiv& dz = (A == 2) ? g.dnz1 : g.dnz;
  const FastDiv& dy = (A == 1) ? g.dny1 : g.dny;
  uint32_t n = (uint32_t)idx;
  uint32_t r = fdiv(n, dz);
  k = (int)(n - r * dz.d);
  uint32_t q = fdiv(r, dy);
  j = (int)(r - q * dy.d);
  i = (int)q;
}
// Generic decomposition for other dims (marching-cubes voxels); 64-bit safe.
__host__ __device__ inline void ijk3(int64_t idx, int hh, int d, int& i, int& j, int& k) {
  i = (int)(idx / d / hh);
  j = (int)((idx / d) % hh);
  k = (int)(idx % d);
}
// Face array index for axis a (ux: (nx+1,ny,nz), uy: (nx,ny+1,nz), uz: (nx,ny,nz+1)).
__host__ __device__ inline int64_t fid(const Grid& g, int a, int i, int j, int k) {
  if (a == 0) return ((int64_t)i * g.ny + j) * g.nz + k;
  if (a == 1) return ((int64_t)i * (g.ny + 1) + j) * g.nz + k;
  return ((int64_t)i * g.ny + j) * (g.nz + 1) + k;
}

// Device-side scalars of an iterative solve; lives in device memory so the
// recurrence runs without host round trips.
struct SolveScalars {
  double rz, rz_prev, pq, rr, bb, alpha, beta, rr_true;
  double al[2];  // alpha of update k in slot k&1 (x may take it one iteration later)
  int iter, done, converged, nonfinite, refinements, rec_conv, fold_pending;
  int last_true;  // MIXED: update index whose residual the last executed true-residual pass replaced (-1 none)
  int done_k;     // single-kernel iteration: the k of the launch that set done
  float alh[2];   // single-kernel iteration: the alpha launch k applied, in slot k&1
};

struct KernelStat {
  std::string name;
  int64_t count = 0;
  double ms = 0;
};

struct PendingEvent {
  int stat;
  hipEvent_t a, b;
};

struct Ctx {
  mpf_config cfg;
  Grid g;
  hipStream_t stream = nullptr;
  std::string err;
  std::string note;  // why the last solve stopped without converging (not an error)
  int device = 0;

  int64_t P = 0, cap = 0;
  int cur = 0;  // which particle buffer is current
  // particles, double-buffered for the sort
  f3* pos[2] = {};
  f3* vel[2] = {};
  int8_t* flag[2] = {};
  float* mass[2] = {};
  float* temp[2] = {};
  float* lheat[2] = {};
  float* solub[2] = {};
  float* gas[2] = {};
  uint32_t* key = nullptr;    // per particle: unsorted keys
  uint32_t* perm = nullptr;   // sorted slot -> old index (gridIndex)
  uint32_t* ghash = nullptr;  // sorted keys (gridHash)
  bool sorted = false;
  bool user_cstart = false, user_cend = false;  // cell lists set by the caller since the last sort
  // the cell lists and particle order come from this library's last sort and
  // no particle moved or was added since (a new sort would be the identity)
  bool own_sort = false;

  // cells
  uint32_t* count = nullptr;  // per-cell counts, then scatter cursors
  uint32_t* cstart = nullptr;
  uint32_t* cend = nullptr;
  uint8_t* cmask = nullptr;
  uint8_t* bmask = nullptr;  // per 4x4x4 brick: OR of cmask, then the same over the 3x3x3 bricks around it
  int8_t* mark = nullptr;
  float *lsfluid = nullptr, *lsair = nullptr, *lsmerge = nullptr;
  float *phig[3] = {};
  float *phifluid = nullptr, *phiair = nullptr, *phisolid = nullptr;
  float* ctmp = nullptr;  // scratch cell field for two-buffer sweeps
  float* sf = nullptr;    // surface tension force per cell
  uint8_t* sflag = nullptr;  // per-tile change flags: 2 x sweep_tiles (eikonal passes), then face_tiles of
                             // each axis (first extrapolation pass)
  void* div = nullptr;    // float, or double under MPF_FP64
  void* press = nullptr;  // float under MPF_FP32, else double
  float *tp = nullptr, *tpchange = nullptr, *tsave = nullptr;
  // iterative solve work vectors (sized for double)
  void *vx = nullptr, *vr = nullptr, *vp = nullptr, *vq = nullptr, *vp2 = nullptr;
  void *vr2 = nullptr, *vq2 = nullptr;  // second r and q of the single-kernel iteration
  void* vp3 = nullptr;                   // third p: the single-kernel iteration keeps p_k-2
  uint8_t* nbm = nullptr;  // per-cell neighbour mask of the current solve
  double* dslow = nullptr;  // 1/diag of the current solve's slow rows (flags-dependent)
  double* vx64 = nullptr;
  double* partials = nullptr;
  int npart = 0;
  SolveScalars* scal = nullptr;
  SolveScalars* scal_host = nullptr;  // pinned
  uint32_t* counters = nullptr;       // [0] reduction arrival, [1..63] misc counters, then spread counters
  uint32_t* counters_host = nullptr;  // pinned
  double* red_host = nullptr;         // pinned, 16 doubles: reduced totals read by the host

  // faces
  float* wu[3] = {};
  float* au[3] = {};
  float* wuo[3] = {};
  float* auo[3] = {};
  float* su[3] = {};
  float* ftmp[3] = {};
  float* p2g_parts = nullptr;  // per-tile face sums of the liquid P2G

  // scan scratch
  uint32_t* scan_blk = nullptr;
  int64_t scan_blk_n = 0;
  uint32_t* ctmp_u32 = nullptr;  // G
  // marching cubes
  uint8_t* mc_case = nullptr;
  uint32_t* mc_edge = nullptr;  // 3G flags -> vertex ids
  uint32_t* mc_tcount = nullptr;  // per voxel triangle offsets
  float* mc_verts = nullptr;
  float* mc_normals = nullptr;   // per vertex (also the smoothing's second buffer)
  uint32_t* mc_vedge = nullptr;  // vertex -> edge id
  uint32_t* mc_tris = nullptr;
  int64_t mc_nv = 0, mc_nt = 0, mc_nvox = 0;
  // particle-source marching cubes (cfg.mc_source != MPF_MC_LEVELSET): the
  // 2N hash grid, its cell lists and the node field
  Grid g2{};
  uint32_t *mc_count = nullptr, *mc_cstart = nullptr, *mc_cend = nullptr;  // g2.G each
  float* mc_field = nullptr;  // (2nx+1)(2ny+1)(2nz+1) nodes
  bool mc_sorted = false;     // particles are in the order of the last 2N sort
  bool mc_field_valid = false;
  // adjoint scratch (allocated on first use): 6 face cotangent arrays in and
  // out, 2 cell fields, per-particle cotangents
  double* adj_fin[6] = {};
  double* adj_fout[6] = {};
  double* adj_cell[2] = {};
  double* adj_part[3] = {};  // vel out / vel in / pos, 3P each
  bool adj_ready = false;
  // nucleation seeds (ring position of the next refresh)
  int32_t* seeds = nullptr;
  int32_t nseeds = 0;
  int32_t seed_ring = 0;
  std::vector<int32_t> seeds_host;
  // pour source (positions, velocities, particle type)
  f3* pour_pos = nullptr;
  f3* pour_vel = nullptr;
  int64_t npour = 0;
  int8_t pour_flag = MPF_TYPEFLUID;

  // timing
  bool timing = false;
  std::vector<KernelStat> stats;
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> event_pool;

  int stat_index(const char* name);
  void flush_timing();
  hipEvent_t get_event();
};

// ---- launch helper: optional HIP-event timing around every launch ----------
bool launch_begin(Ctx& c, const char* name, PendingEvent& pe);
void launch_end(Ctx& c, PendingEvent& pe, bool timed);

template <typename... KArgs, typename... Args>
inline void launch(Ctx& c, const char* name, dim3 grid, dim3 block, size_t smem, void (*k)(KArgs...),
                   Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  PendingEvent pe;
  bool timed = launch_begin(c, name, pe);
  hipLaunchKernelGGL(k, grid, block, smem, c.stream, static_cast<KArgs>(args)...);
  launch_end(c, pe, timed);
}

inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- stage entry points implemented in the .hip files ----------------------
// sort.hip
void run_hash_sort(Ctx& c);
void run_cell_mask(Ctx& c);
void run_bricks(Ctx& c);
int64_t run_check_cells(Ctx& c);  // cells whose list does not index the particles; -1 on a HIP error
int64_t run_scan_u32(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n, bool want_total);
// particles.hip
void run_apply_forces(Ctx& c);
void run_g2p(Ctx& c);
void run_advect(Ctx& c);
void run_heat_g2p(Ctx& c);
int run_phase_change(Ctx& c, int frame, mpf_phase_info* info);
int run_phase_change_with_delete(Ctx& c, int frame, mpf_phase_info* info, int64_t* deleted);
int run_delete_particles(Ctx& c, int64_t* deleted, int64_t* fluid_deleted);
int run_pour(Ctx& c, int frame, int64_t* added);
// grid.hip
void run_levelset(Ctx& c);
void run_p2g(Ctx& c);
void run_extrapolate_bc(Ctx& c, bool recompute_phi);
void run_divergence(Ctx& c);
void run_subgrad(Ctx& c);
void run_heat_grid_pre(Ctx& c);
void run_heat_grid_post(Ctx& c);
void run_heat_jacobi(Ctx& c, int iters);
int run_marching_cubes(Ctx& c, mpf_mc_info* info);
void run_hash_sort_mc(Ctx& c);
// bubble.hip
void run_mark_solo_air(Ctx& c);
void run_surface_tension(Ctx& c);
void run_drag(Ctx& c);
void run_solo_air_sph(Ctx& c);
void run_correct_pos(Ctx& c);
void run_correct_bubble_pos(Ctx& c);
void sph_constants(float h, float& poly6, float& spiky, float& lap);
// solid.hip
int run_solid_motion(Ctx& c);
// io.hip
int run_write_mesh(Ctx& c, const char* path, int format, const char* object_name);
void run_solid_phi_and_velocity(Ctx& c);
int64_t p2g_parts_floats(const Grid& g);
// Device counters: 64 single counters, then SPREAD counters 64 bytes apart
// (a total that many blocks add to, split over cache lines)
constexpr int NSPREAD = 64;
constexpr int SPREAD_STRIDE = 16;
constexpr int NCOUNTERS = 64 + NSPREAD * SPREAD_STRIDE;
// 1x4x16 cell tiles of the eikonal passes, one wave's cells each (change
// flags per tile and pass)
constexpr int SW_J = 4, SW_K = 16;
inline int64_t sweep_tiles(const Grid& g) {
  return (int64_t)g.nx * ((g.ny + SW_J - 1) / SW_J) * ((g.nz + SW_K - 1) / SW_K);
}
// the same tiles over the faces of axis a (extrapolation passes)
inline int64_t face_tiles(const Grid& g, int a) {
  return (int64_t)(g.nx + (a == 0)) * ((g.ny + (a == 1) + SW_J - 1) / SW_J) * ((g.nz + (a == 2) + SW_K - 1) / SW_K);
}
// 4x4x4 cell bricks of the level-set search summary
inline int64_t nbricks(const Grid& g) {
  return (int64_t)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
}
// pcg.hip
int run_pressure_solve(Ctx& c, mpf_solve_info* info);
int run_heat_cg(Ctx& c, mpf_solve_info* info);
int run_adjoint_solve(Ctx& c, const double* rhs, double* sol, mpf_solve_info* info);
// adjoint.hip
bool adj_alloc_all(Ctx& c);
void adj_free(Ctx& c);
int run_project_adjoint(Ctx& c, const double* const* gout, double* const* gin, int where, mpf_solve_info* info);
int run_extrapolate_adjoint(Ctx& c, const double* const* gout, double* const* gin, int where);
int run_p2g_adjoint(Ctx& c, const double* const* gfaces, double* gvel, double* gpos, int where);
int run_g2p_adjoint(Ctx& c, const double* gvel_out, double* gvel_in, double* const* gu, double* const* guo,
                    double* gpos, int where);
int run_advect_adjoint(Ctx& c, const double* gpos_out, const double* gvel_out, double* gpos_in, double* gvel_in,
                       int where, int64_t* in_band);

bool hip_ok(Ctx& c, hipError_t e, const char* what);
extern thread_local std::string g_noctx_err;

}  // namespace mpf
