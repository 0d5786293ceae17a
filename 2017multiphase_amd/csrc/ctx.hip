// Host side of the C ABI (include/mpflip.h): context lifetime, device memory,
// data movement, the per-stage entry points and the composite step.
//
// The context plays the role of `class cspray` (spray.h:18-389): it owns every
// device buffer (initmem_bubble, spray.cu:25-218) and runs each stage on one
// HIP stream.  Unlike the reference, nothing synchronises after each launch;
// the host waits only for solver polls and count readbacks.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "mpf_internal.h"

namespace mpf {

bool hip_ok(Ctx& c, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  c.err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

int Ctx::stat_index(const char* name) {
  for (size_t i = 0; i < stats.size(); ++i)
    if (stats[i].name == name) return (int)i;
  stats.push_back(KernelStat{name, 0, 0.0});
  return (int)stats.size() - 1;
}

hipEvent_t Ctx::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

void Ctx::flush_timing() {
  if (pending.empty()) return;
  hipStreamSynchronize(stream);
  for (auto& pe : pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      stats[pe.stat].ms += ms;
      stats[pe.stat].count += 1;
    }
    event_pool.push_back(pe.a);
    event_pool.push_back(pe.b);
  }
  pending.clear();
}

bool launch_begin(Ctx& c, const char* name, PendingEvent& pe) {
  if (!c.timing) return false;
  pe.stat = c.stat_index(name);
  pe.a = c.get_event();
  pe.b = c.get_event();
  hipEventRecord(pe.a, c.stream);
  return true;
}

void launch_end(Ctx& c, PendingEvent& pe, bool timed) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess && c.err.empty()) c.err = std::string("kernel launch: ") + hipGetErrorString(e);
  if (!timed) return;
  hipEventRecord(pe.b, c.stream);
  c.pending.push_back(pe);
  if (c.pending.size() > 4096) c.flush_timing();
}

}  // namespace mpf

using namespace mpf;

struct mpf_ctx {
  Ctx c;
};

namespace mpf {
// message of the last failing call made without a context (mpf_create,
// mpf_hash_sort_noalloc), returned by mpf_last_error(NULL)
thread_local std::string g_noctx_err;
}  // namespace mpf

namespace {

template <class T>
bool dalloc(Ctx& c, T*& p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) {
    c.err = std::string("hipMalloc failed (") + hipGetErrorString(e) + ")";
    p = nullptr;
    return false;
  }
  p = (T*)q;
  return true;
}

template <class T>
void dfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

void free_all(Ctx& c) {
  for (int b = 0; b < 2; ++b) {
    dfree(c.pos[b]); dfree(c.vel[b]); dfree(c.flag[b]); dfree(c.mass[b]);
    dfree(c.temp[b]); dfree(c.lheat[b]); dfree(c.solub[b]); dfree(c.gas[b]);
  }
  dfree(c.key); dfree(c.perm); dfree(c.ghash);
  dfree(c.count); dfree(c.cstart); dfree(c.cend); dfree(c.cmask); dfree(c.bmask); dfree(c.sflag); dfree(c.mark);
  dfree(c.lsfluid); dfree(c.lsair); dfree(c.lsmerge);
  for (int a = 0; a < 3; ++a) dfree(c.phig[a]);
  dfree(c.phifluid); dfree(c.phiair); dfree(c.phisolid); dfree(c.ctmp); dfree(c.sf);
  dfree(c.div); dfree(c.press); dfree(c.tp); dfree(c.tpchange); dfree(c.tsave);
  dfree(c.vx); dfree(c.vr); dfree(c.vp); dfree(c.vq); dfree(c.vp2); dfree(c.vr2); dfree(c.vq2); dfree(c.vp3); dfree(c.nbm); dfree(c.dslow);
  dfree(c.partials); dfree(c.scal); dfree(c.counters);
  for (int a = 0; a < 3; ++a) {
    dfree(c.wu[a]); dfree(c.au[a]); dfree(c.wuo[a]); dfree(c.auo[a]); dfree(c.su[a]); dfree(c.ftmp[a]);
  }
  dfree(c.scan_blk); dfree(c.ctmp_u32); dfree(c.p2g_parts);
  dfree(c.mc_case); dfree(c.mc_edge); dfree(c.mc_tcount); dfree(c.mc_verts); dfree(c.mc_tris);
  dfree(c.mc_normals); dfree(c.mc_vedge);
  dfree(c.mc_count); dfree(c.mc_cstart); dfree(c.mc_cend); dfree(c.mc_field);
  dfree(c.seeds);
  dfree(c.pour_pos);
  dfree(c.pour_vel);
  adj_free(c);
  if (c.scal_host) hipHostFree(c.scal_host);
  if (c.counters_host) hipHostFree(c.counters_host);
  if (c.red_host) hipHostFree(c.red_host);
  c.scal_host = nullptr;
  c.counters_host = nullptr;
  c.red_host = nullptr;
  for (auto& pe : c.pending) { hipEventDestroy(pe.a); hipEventDestroy(pe.b); }
  for (auto e : c.event_pool) hipEventDestroy(e);
  c.pending.clear();
  c.event_pool.clear();
  if (c.stream) hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// Nodes and voxels of the marching-cubes lattice of the configured source.
void mc_lattice(const Ctx& c, int64_t& nodes, int64_t& voxels) {
  if (c.cfg.mc_source == MPF_MC_LEVELSET) {
    nodes = c.g.G;
    voxels = (int64_t)(c.g.nx - 1) * (c.g.ny - 1) * (c.g.nz - 1);
  } else {
    nodes = (int64_t)(c.g2.nx + 1) * (c.g2.ny + 1) * (c.g2.nz + 1);
    voxels = c.g2.G;
  }
}

bool alloc_all(Ctx& c) {
  const int64_t G = c.g.G, cap = c.cap;
  int64_t MN = 0, V = 0;
  mc_lattice(c, MN, V);
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  bool ok = true;
  for (int b = 0; b < 2 && ok; ++b) {
    ok = ok && dalloc(c, c.pos[b], cap) && dalloc(c, c.vel[b], cap) && dalloc(c, c.flag[b], cap) &&
         dalloc(c, c.mass[b], cap) && dalloc(c, c.temp[b], cap) && dalloc(c, c.lheat[b], cap) &&
         dalloc(c, c.solub[b], cap) && dalloc(c, c.gas[b], cap);
  }
  ok = ok && dalloc(c, c.key, cap) && dalloc(c, c.perm, cap) && dalloc(c, c.ghash, cap);
  ok = ok && dalloc(c, c.count, G) && dalloc(c, c.cstart, G) && dalloc(c, c.cend, G) && dalloc(c, c.cmask, G) &&
       dalloc(c, c.mark, G) && dalloc(c, c.lsfluid, G) && dalloc(c, c.lsair, G) && dalloc(c, c.lsmerge, G) &&
       dalloc(c, c.bmask, 2 * nbricks(c.g)) &&
       dalloc(c, c.sflag, 2 * sweep_tiles(c.g) + face_tiles(c.g, 0) + face_tiles(c.g, 1) + face_tiles(c.g, 2));
  for (int a = 0; a < 3; ++a) ok = ok && dalloc(c, c.phig[a], G);
  ok = ok && dalloc(c, c.phifluid, G) && dalloc(c, c.phiair, G) && dalloc(c, c.phisolid, G) && dalloc(c, c.ctmp, G) &&
       dalloc(c, c.sf, G);
  double* dd = nullptr;
  ok = ok && dalloc(c, dd, G);
  c.div = dd;
  ok = ok && dalloc(c, dd, G);
  c.press = dd;
  ok = ok && dalloc(c, c.tp, G) && dalloc(c, c.tpchange, G) && dalloc(c, c.tsave, G);
  ok = ok && dalloc(c, dd, G); c.vx = dd;
  ok = ok && dalloc(c, dd, G); c.vr = dd;
  ok = ok && dalloc(c, dd, G); c.vp = dd;
  ok = ok && dalloc(c, dd, G); c.vq = dd;
  ok = ok && dalloc(c, dd, G); c.vp2 = dd;
  ok = ok && dalloc(c, dd, G); c.vr2 = dd;
  ok = ok && dalloc(c, dd, G); c.vq2 = dd;
  ok = ok && dalloc(c, dd, G); c.vp3 = dd;
  ok = ok && dalloc(c, c.nbm, G) && dalloc(c, c.dslow, G);
  c.npart = 1024 * 48;  // solver partial sums (pcg.hip lays out regions of 1024-block slots)
  ok = ok && dalloc(c, c.partials, c.npart) && dalloc(c, c.scal, 1) && dalloc(c, c.counters, NCOUNTERS);
  for (int a = 0; a < 3 && ok; ++a)
    ok = ok && dalloc(c, c.wu[a], fc[a]) && dalloc(c, c.au[a], fc[a]) && dalloc(c, c.wuo[a], fc[a]) &&
         dalloc(c, c.auo[a], fc[a]) && dalloc(c, c.su[a], fc[a]) && dalloc(c, c.ftmp[a], fc[a]);
  int64_t maxscan = std::max<int64_t>(3 * G, std::max<int64_t>(V, cap));
  if (c.cfg.mc_enabled) maxscan = std::max<int64_t>(maxscan, std::max<int64_t>(3 * MN, c.g2.G));
  c.scan_blk_n = maxscan / 4096 + 4;
  ok = ok && dalloc(c, c.scan_blk, c.scan_blk_n) && dalloc(c, c.ctmp_u32, G);
  ok = ok && dalloc(c, c.p2g_parts, (size_t)p2g_parts_floats(c.g));
  if (ok && c.cfg.mc_enabled) {
    ok = ok && dalloc(c, c.mc_case, V > 0 ? V : 1) && dalloc(c, c.mc_edge, 6 * MN) &&
         dalloc(c, c.mc_tcount, 2 * (V > 0 ? V : 1)) && dalloc(c, c.mc_verts, 3 * c.cfg.mc_max_verts) &&
         dalloc(c, c.mc_tris, 3 * c.cfg.mc_max_tris) && dalloc(c, c.mc_normals, 3 * c.cfg.mc_max_verts) &&
         dalloc(c, c.mc_vedge, c.cfg.mc_max_verts);
    if (c.cfg.mc_source != MPF_MC_LEVELSET)
      ok = ok && dalloc(c, c.mc_count, c.g2.G) && dalloc(c, c.mc_cstart, c.g2.G) && dalloc(c, c.mc_cend, c.g2.G) &&
           dalloc(c, c.mc_field, MN);
  }
  if (ok && c.cfg.adjoint_enabled) ok = adj_alloc_all(c);
  if (!ok) {
    if (c.err.empty()) c.err = "hipMalloc failed (adjoint scratch)";
    return false;
  }
  if (hipHostMalloc((void**)&c.scal_host, sizeof(SolveScalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c.counters_host, NCOUNTERS * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c.red_host, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    c.err = "hipHostMalloc failed";
    return false;
  }
  hipMemsetAsync(c.counters, 0, NCOUNTERS * 4, c.stream);
  hipMemsetAsync(c.scal, 0, sizeof(SolveScalars), c.stream);
  // zero-initialise the fields a caller may read before they are computed
  hipMemsetAsync(c.mark, VACUUM, (size_t)G, c.stream);
  hipMemsetAsync(c.cmask, 0, (size_t)G, c.stream);
  hipMemsetAsync(c.cstart, 0xff, (size_t)G * 4, c.stream);
  hipMemsetAsync(c.cend, 0xff, (size_t)G * 4, c.stream);
  for (float* f : {c.lsfluid, c.lsair, c.lsmerge, c.phig[0], c.phig[1], c.phig[2], c.phifluid, c.phiair, c.phisolid,
                   c.tp, c.tpchange, c.tsave, c.sf})
    hipMemsetAsync(f, 0, (size_t)G * 4, c.stream);
  hipMemsetAsync(c.div, 0, (size_t)G * 8, c.stream);
  hipMemsetAsync(c.press, 0, (size_t)G * 8, c.stream);
  for (int a = 0; a < 3; ++a)
    for (float* f : {c.wu[a], c.au[a], c.wuo[a], c.auo[a], c.su[a]}) hipMemsetAsync(f, 0, (size_t)fc[a] * 4, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "initial clear");
}

mpf_status fail(Ctx& c, mpf_status s, const std::string& msg) {
  c.err = msg;
  return s;
}

mpf_status check_launch(Ctx& c) {
  if (!c.err.empty()) return MPF_HIP_ERROR;
  return MPF_OK;
}

// Element pointer, count and size of a field.
bool field_ref(Ctx& c, int f, void** ptr, int64_t* n, int* es) {
  const int64_t G = c.g.G;
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  *es = 4;
  *n = G;
  switch (f) {
    case MPF_F_MARK: *ptr = c.mark; *es = 1; return true;
    case MPF_F_LSFLUID: *ptr = c.lsfluid; return true;
    case MPF_F_LSAIR: *ptr = c.lsair; return true;
    case MPF_F_LSMERGE: *ptr = c.lsmerge; return true;
    case MPF_F_PHIGRAD_X: *ptr = c.phig[0]; return true;
    case MPF_F_PHIGRAD_Y: *ptr = c.phig[1]; return true;
    case MPF_F_PHIGRAD_Z: *ptr = c.phig[2]; return true;
    case MPF_F_PHIFLUID: *ptr = c.phifluid; return true;
    case MPF_F_PHIAIR: *ptr = c.phiair; return true;
    case MPF_F_PHISOLID: *ptr = c.phisolid; return true;
    case MPF_F_DIV: *ptr = c.div; *es = c.cfg.precision == MPF_FP64 ? 8 : 4; return true;
    case MPF_F_PRESSURE: *ptr = c.press; *es = c.cfg.precision == MPF_FP32 ? 4 : 8; return true;
    case MPF_F_TEMPERATURE: *ptr = c.tp; return true;
    case MPF_F_TEMPERATURE_CHANGE: *ptr = c.tpchange; return true;
    case MPF_F_CELL_START: *ptr = c.cstart; return true;
    case MPF_F_CELL_END: *ptr = c.cend; return true;
    case MPF_F_GRID_HASH: *ptr = c.ghash; *n = c.P; return true;
    case MPF_F_GRID_INDEX: *ptr = c.perm; *n = c.P; return true;
    case MPF_F_SURFACE_TENSION: *ptr = c.sf; return true;
    default: break;
  }
  float** groups[5] = {c.wu, c.au, c.wuo, c.auo, c.su};
  if (f >= MPF_F_WATER_UX && f <= MPF_F_SOLID_UZ) {
    int q = f - MPF_F_WATER_UX;
    *ptr = groups[q / 3][q % 3];
    *n = fc[q % 3];
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

int32_t mpf_abi_version(void) { return MPF_ABI_VERSION; }

mpf_status mpf_config_default(mpf_config* cfg, int32_t nx, int32_t ny, int32_t nz) {
  if (!cfg || nx < 4 || ny < 4 || nz < 4) return MPF_INVALID_ARG;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nx = nx; cfg->ny = ny; cfg->nz = nz;
  cfg->dx = 1.0f / 64;                       // spray.cpp:78
  cfg->samplespace = 0.5f * cfg->dx;         // spray.cpp:79
  cfg->dt = 0.003f;                          // spray.cpp:575
  cfg->gravity[2] = -9.8f;                   // spray.cpp:77
  cfg->m0 = (1.0f * 1000.0f) / powf(1.0f / cfg->samplespace, 3.0f);  // spray.cpp:581
  cfg->airm0 = cfg->m0;                      // spray.cpp:582
  cfg->particle_capacity = (int64_t)nx * ny * nz * 10;
  cfg->buoyancy_kick = 0;
  cfg->bubble_max_vel = 6.5f;                // spray.cpp:66
  cfg->buoyancy_rate_air = 1.25f;            // spray.cpp:71
  cfg->buoyancy_rate_solo = 1.05f;
  cfg->precision = MPF_MIXED;
  cfg->solver = MPF_SOLVER_JACOBI_PCG;
  cfg->op = MPF_OPERATOR_VARDENS;
  cfg->rho_liquid = 1000.0f;                 // spray.cpp:579
  cfg->rho_air = 1.0f;
  cfg->tol = 1e-6;
  cfg->max_iters = 4000;
  cfg->ref_max_iters = 800;                  // spray.cu:1642-1649
  cfg->refine_every = 20;
  cfg->check_every = 16;
  cfg->flip_alpha = 1.0f;                    // bubblesim G2P is pure FLIP
  cfg->integrator = MPF_ADVECT_EULER_COLLIDE;
  cfg->fix_solids = 1;
  cfg->bounce_vel = 1.0f;                    // spray.cpp:55
  cfg->bounce_pos = 0.0f;
  cfg->heat_enabled = 1;
  cfg->heat_solver = MPF_HEAT_CG;
  cfg->heat_jacobi_iters = 10;
  cfg->heat_max_iters = 3800;                // spray.cu:1803
  cfg->heat_alpha[MPF_TYPEFLUID] = 0.02f;    // spray.cpp:35-36, spray.cu:3076-3090
  cfg->heat_alpha[MPF_TYPEAIR] = 0.008f;
  cfg->heat_alpha[MPF_TYPEBOUNDARY] = 0.0f;
  cfg->heat_alpha[MPF_TYPEVACUUM] = 0.0001f;
  cfg->heat_alpha[MPF_TYPEAIRSOLO] = 0.008f;
  cfg->heat_alpha[MPF_TYPESOLID] = 0.2f;
  cfg->default_temperature = 293.15f;        // spray.cpp:37
  cfg->heat_add_bottom = 1;
  cfg->heat_increase_bottom = 1.5f;
  cfg->heat_bottom_clamp = 378.0f;           // spray_k.cu:4502
  cfg->heat_extend_boundary = 0;
  cfg->heat_vacuum_neumann = 0;
  cfg->heat_melt_and_boil = 0;
  cfg->alpha_temp_trans = 0.0f;              // spray.cpp:54
  cfg->melt_enabled = 0;
  cfg->freeze_enabled = 0;
  cfg->gas_enabled = 1;
  cfg->gas_every = 2;                        // spray.cpp:1159
  cfg->melting_point = 273.15f;              // spray.cpp:38
  cfg->boiling_point = 373.15f;
  cfg->liquid_heat_th = 10.0f;               // spray.cpp:52
  cfg->temperature0 = 273.15f;
  cfg->dissolve_gas_rate = 1.0f;             // spray.cpp:57
  cfg->vapor_gen_rate = 0.01f;               // spray.cpp:58
  cfg->mc_enabled = 1;
  int64_t G = (int64_t)nx * ny * nz;
  cfg->mc_max_verts = G > (1 << 20) ? G : (1 << 20);
  cfg->mc_max_tris = cfg->mc_max_verts;
  cfg->seed = 0;
  cfg->adjoint_enabled = 0;
  cfg->delete_enabled = 0;
  cfg->delete_vacuum_sweeps = 20;            // spray.cu:2205-2206
  cfg->pour_every = 4;                       // spray.cu:3014
  cfg->pour_pos_rand = 0.1f;                 // spray.cpp:55
  cfg->pour_vel_rand = 0.0f;
  cfg->pour_temperature = 293.15f;           // spray.cpp:54
  cfg->seed_update_every = 0;
  cfg->mc_source = MPF_MC_LEVELSET;
  cfg->mc_gas_airsolo = 1;
  cfg->mc_every = 1;
  cfg->mc_smooth_iters = 5;                  // m_bSmoothMC, smoothIterTimes (spray.cpp:62, 610)
  cfg->solo_air_enabled = 1;
  cfg->solo_tol_in = -1.45f;                 // spray_k.cu:4245
  cfg->solo_tol_out = -0.5f;
  cfg->sph_viscosity = 1.8f;                 // spray.cpp:40
  cfg->surface_tension_sigma = 1.0f;         // spray.cpp:34
  cfg->drag_enabled = 1;
  cfg->drag_solo = 10.0f;                    // spray.cpp:39
  cfg->drag_grid = 0.05f;
  cfg->drag_particle_update = 0;
  cfg->correct_position = 1;                 // spray.cpp:50
  cfg->correction_spring = 500000.0f;        // spray.cpp:617-618
  cfg->correction_radius = 0.5f;
  cfg->correct_bubble_air_ls = 0;
  cfg->solid_bounce = 0.5f;                  // SolidbounceParam, spray.cpp:52
  cfg->rigid_lock_y = 0;
  cfg->pcg_recompute_q = 0;
  return MPF_OK;
}

mpf_status mpf_create(const mpf_config* cfg, mpf_ctx** out) {
  if (!cfg || !out) {
    g_noctx_err = "null argument";
    return MPF_INVALID_ARG;
  }
  *out = nullptr;
  if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4 || !(cfg->dx > 0) || cfg->particle_capacity < 0 ||
      (int64_t)cfg->nx * cfg->ny * cfg->nz * 3 > 0xffffffffll || cfg->particle_capacity > 0xfffffff0ll ||
      cfg->precision < 0 || cfg->precision > 2 || cfg->solver < 0 || cfg->solver > 1 || cfg->op < 0 || cfg->op > 1 ||
      (cfg->op == MPF_OPERATOR_VARDENS && !(cfg->rho_liquid > 0 && cfg->rho_air > 0)) ||
      (cfg->heat_enabled && cfg->heat_solver == MPF_HEAT_CG && cfg->heat_alpha[0] <= 0)) {
    g_noctx_err = "invalid configuration";
    return MPF_INVALID_ARG;
  }
  if (cfg->integrator < MPF_ADVECT_EULER_COLLIDE || cfg->integrator > MPF_ADVECT_FREEZING) {
    g_noctx_err = "invalid configuration (integrator)";
    return MPF_INVALID_ARG;
  }
  if (cfg->mc_source < MPF_MC_LEVELSET || cfg->mc_source > MPF_MC_SOLID || cfg->mc_every < 1) {
    g_noctx_err = "invalid marching cubes configuration (mc_source, mc_every)";
    return MPF_INVALID_ARG;
  }
  if (cfg->mc_enabled && cfg->mc_source != MPF_MC_LEVELSET &&
      (int64_t)(2 * cfg->nx + 1) * (2 * cfg->ny + 1) * (2 * cfg->nz + 1) * 3 > 0xffffffffll) {
    g_noctx_err = "grid too large for the 2N marching cubes lattice (3 edges per node must fit 32 bits)";
    return MPF_INVALID_ARG;
  }
  mpf_ctx* h = new (std::nothrow) mpf_ctx;
  if (!h) return MPF_CAPACITY;
  Ctx& c = h->c;
  c.cfg = *cfg;
  if (c.cfg.check_every <= 0) c.cfg.check_every = 16;
  Grid& g = c.g;
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.G = (int64_t)g.nx * g.ny * g.nz;
  g.FX = (int64_t)(g.nx + 1) * g.ny * g.nz;
  g.FY = (int64_t)g.nx * (g.ny + 1) * g.nz;
  g.FZ = (int64_t)g.nx * g.ny * (g.nz + 1);
  g.h = cfg->dx;
  g.gx0 = cfg->gmin[0]; g.gy0 = cfg->gmin[1]; g.gz0 = cfg->gmin[2];
  g.gx1 = cfg->gmin[0] + cfg->dx * (float)g.nx;
  g.gy1 = cfg->gmin[1] + cfg->dx * (float)g.ny;
  g.gz1 = cfg->gmin[2] + cfg->dx * (float)g.nz;
  g.dnz = make_fastdiv((uint32_t)g.nz);
  g.dny = make_fastdiv((uint32_t)g.ny);
  g.dnz1 = make_fastdiv((uint32_t)g.nz + 1);
  g.dny1 = make_fastdiv((uint32_t)g.ny + 1);
  // 2N hash grid of preMC (calcHashD_MC, spray_k.cu:1038-1042): cell size
  // cellsize.x / NXMC * NX evaluated in float
  Grid& g2 = c.g2;
  g2.nx = 2 * g.nx; g2.ny = 2 * g.ny; g2.nz = 2 * g.nz;
  g2.G = (int64_t)g2.nx * g2.ny * g2.nz;
  g2.h = (cfg->dx / (float)g2.nx) * (float)g.nx;
  g2.gx0 = g.gx0; g2.gy0 = g.gy0; g2.gz0 = g.gz0;
  g2.dnz = make_fastdiv((uint32_t)g2.nz);
  g2.dny = make_fastdiv((uint32_t)g2.ny);
  g2.dnz1 = make_fastdiv((uint32_t)g2.nz + 1);
  g2.dny1 = make_fastdiv((uint32_t)g2.ny + 1);
  c.cap = cfg->particle_capacity;
  hipGetDevice(&c.device);
  if (!hip_ok(c, hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "stream")) {
    g_noctx_err = c.err;
    delete h;
    return MPF_HIP_ERROR;
  }
  if (!alloc_all(c)) {
    g_noctx_err = c.err;
    free_all(c);
    delete h;
    return MPF_CAPACITY;
  }
  *out = h;
  return MPF_OK;
}

mpf_status mpf_destroy(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  hipStreamSynchronize(ctx->c.stream);
  free_all(ctx->c);
  delete ctx;
  return MPF_OK;
}

const char* mpf_last_error(mpf_ctx* ctx) {
  if (!ctx) return g_noctx_err.c_str();
  return ctx->c.err.c_str();
}

void* mpf_stream(mpf_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

mpf_status mpf_synchronize(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  return hip_ok(ctx->c, hipStreamSynchronize(ctx->c.stream), "synchronize") ? MPF_OK : MPF_HIP_ERROR;
}

int64_t mpf_particle_count(mpf_ctx* ctx) { return ctx ? ctx->c.P : -1; }

mpf_status mpf_set_particles(mpf_ctx* ctx, int64_t n, const float* pos, const float* vel, const int8_t* flag,
                             const float* mass, const float* T, const float* lheat, const float* solub,
                             const float* gas, int32_t where) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (n < 0) return fail(c, MPF_INVALID_ARG, "negative particle count");
  if (n > c.cap) return fail(c, MPF_CAPACITY, "particle count exceeds capacity");
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const int b = c.cur;
  auto put = [&](void* dst, const void* src, size_t bytes, int fillbyte) {
    if (src) hipMemcpyAsync(dst, src, bytes, kind, c.stream);
    else hipMemsetAsync(dst, fillbyte, bytes, c.stream);
  };
  put(c.pos[b], pos, (size_t)n * 12, 0);
  put(c.vel[b], vel, (size_t)n * 12, 0);
  put(c.flag[b], flag, (size_t)n, MPF_TYPEFLUID);
  put(c.mass[b], mass, (size_t)n * 4, 0);
  put(c.temp[b], T, (size_t)n * 4, 0);
  put(c.lheat[b], lheat, (size_t)n * 4, 0);
  put(c.solub[b], solub, (size_t)n * 4, 0);
  put(c.gas[b], gas, (size_t)n * 4, 0);
  c.P = n;
  c.sorted = false;
  c.own_sort = false;
  c.mc_sorted = false;
  c.user_cstart = c.user_cend = false;
  return hip_ok(c, hipStreamSynchronize(c.stream), "set_particles") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_get_particles(mpf_ctx* ctx, int64_t* n, float* pos, float* vel, int8_t* flag, float* mass, float* T,
                             float* lheat, float* solub, float* gas, int32_t where) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  const int b = c.cur;
  const size_t P = (size_t)c.P;
  if (n) *n = c.P;
  auto get = [&](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) hipMemcpyAsync(dst, src, bytes, kind, c.stream);
  };
  get(pos, c.pos[b], P * 12);
  get(vel, c.vel[b], P * 12);
  get(flag, c.flag[b], P);
  get(mass, c.mass[b], P * 4);
  get(T, c.temp[b], P * 4);
  get(lheat, c.lheat[b], P * 4);
  get(solub, c.solub[b], P * 4);
  get(gas, c.gas[b], P * 4);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_particles") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_field_info(mpf_ctx* ctx, int32_t field, int64_t* count, int32_t* elem_bytes) {
  if (!ctx || !count || !elem_bytes) return MPF_INVALID_ARG;
  void* p;
  int64_t n;
  int es;
  if (!field_ref(ctx->c, field, &p, &n, &es)) return fail(ctx->c, MPF_INVALID_ARG, "unknown field");
  *count = n;
  *elem_bytes = es;
  return MPF_OK;
}

mpf_status mpf_get_field(mpf_ctx* ctx, int32_t field, void* dst, int32_t where) {
  if (!ctx || !dst) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  void* p;
  int64_t n;
  int es;
  if (!field_ref(c, field, &p, &n, &es)) return fail(c, MPF_INVALID_ARG, "unknown field");
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (n > 0)
    hipMemcpyAsync(dst, p, (size_t)n * es, where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                   c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_field") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_set_field(mpf_ctx* ctx, int32_t field, const void* src, int32_t where) {
  if (!ctx || !src) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  void* p;
  int64_t n;
  int es;
  if (!field_ref(c, field, &p, &n, &es)) return fail(c, MPF_INVALID_ARG, "unknown field");
  if (field == MPF_F_GRID_HASH || field == MPF_F_GRID_INDEX) return fail(c, MPF_INVALID_ARG, "read-only field");
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (n > 0)
    hipMemcpyAsync(p, src, (size_t)n * es, where == MPF_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                   c.stream);
  if (field == MPF_F_CELL_START || field == MPF_F_CELL_END) {
    // Caller-supplied cell lists (particles assumed in that order) take effect
    // once both arrays have been given and every list indexes the particles.
    c.sorted = false;
    c.own_sort = false;
    if (field == MPF_F_CELL_START) c.user_cstart = true;
    else c.user_cend = true;
    if (!hip_ok(c, hipStreamSynchronize(c.stream), "set_field")) return MPF_HIP_ERROR;
    if (c.user_cstart && c.user_cend) {
      const int64_t bad = run_check_cells(c);
      if (bad < 0) return MPF_HIP_ERROR;
      if (bad > 0)
        return fail(c, MPF_INVALID_ARG,
                    std::to_string(bad) + " cell lists do not index the particle array (need start < end <= P)");
      run_cell_mask(c);
      c.sorted = true;
    }
  }
  return hip_ok(c, hipStreamSynchronize(c.stream), "set_field") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_set_seed_cells(mpf_ctx* ctx, const int32_t* cells, int32_t n) {
  if (!ctx || n < 0 || (n > 0 && !cells)) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  for (int32_t i = 0; i < n; ++i)
    if (cells[i] < 0 || cells[i] >= c.g.G) return fail(c, MPF_INVALID_ARG, "seed cell out of range");
  hipStreamSynchronize(c.stream);
  dfree(c.seeds);
  c.nseeds = 0;
  c.seed_ring = 0;
  c.seeds_host.assign(cells, cells + n);
  if (n == 0) return MPF_OK;
  if (!dalloc(c, c.seeds, (size_t)n)) return MPF_CAPACITY;
  hipMemcpyAsync(c.seeds, cells, (size_t)n * 4, hipMemcpyHostToDevice, c.stream);
  c.nseeds = n;
  return hip_ok(c, hipStreamSynchronize(c.stream), "set_seed_cells") ? MPF_OK : MPF_HIP_ERROR;
}

// Counter hash of the nucleation and pour random numbers (SplitMix64
// finalizer, Steele, Lea, Flood, OOPSLA 2014), host side.
static uint64_t mix64_host(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

mpf_status mpf_update_seed_cells(mpf_ctx* ctx, int32_t frame) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  const int every = c.cfg.seed_update_every;
  if (every <= 0 || c.nseeds <= 0 || frame % every != 0) return MPF_OK;
  const int bound = 3, z = 2;
  if (c.g.nx <= 2 * bound || c.g.ny <= 2 * bound || c.g.nz <= z) return MPF_OK;
  // four replacements per refresh (the reference's `while (cnt++ <= 3)`)
  const uint64_t tag = 0x73656564ull << 32;
  for (int q = 0; q < 4; ++q) {
    const uint64_t hx = mix64_host(c.cfg.seed ^ mix64_host(tag ^ mix64_host((uint64_t)frame ^ mix64_host(2 * q))));
    const uint64_t hy = mix64_host(c.cfg.seed ^ mix64_host(tag ^ mix64_host((uint64_t)frame ^ mix64_host(2 * q + 1))));
    const int x = bound + (int)((hx >> 11) % (uint64_t)(c.g.nx - 2 * bound));
    const int y = bound + (int)((hy >> 11) % (uint64_t)(c.g.ny - 2 * bound));
    c.seeds_host[c.seed_ring] = (int32_t)cid(c.g, x, y, z);
    c.seed_ring = (c.seed_ring + 1) % c.nseeds;
  }
  hipMemcpyAsync(c.seeds, c.seeds_host.data(), (size_t)c.nseeds * 4, hipMemcpyHostToDevice, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "update_seed_cells") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_set_pour_source(mpf_ctx* ctx, int64_t n, const float* pos, const float* vel, int8_t flag) {
  if (!ctx || n < 0 || (n > 0 && (!pos || !vel))) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (flag != MPF_TYPEFLUID && flag != MPF_TYPEAIRSOLO && flag != MPF_TYPEAIR && flag != MPF_TYPESOLID)
    return fail(c, MPF_INVALID_ARG, "pour source type must be a particle type");
  if (n > c.cap) return fail(c, MPF_CAPACITY, "pour source larger than the particle capacity");
  hipStreamSynchronize(c.stream);
  dfree(c.pour_pos);
  dfree(c.pour_vel);
  c.npour = 0;
  if (n == 0) return MPF_OK;
  if (!dalloc(c, c.pour_pos, (size_t)n) || !dalloc(c, c.pour_vel, (size_t)n)) return MPF_CAPACITY;
  hipMemcpyAsync(c.pour_pos, pos, (size_t)n * 12, hipMemcpyHostToDevice, c.stream);
  hipMemcpyAsync(c.pour_vel, vel, (size_t)n * 12, hipMemcpyHostToDevice, c.stream);
  c.npour = n;
  c.pour_flag = flag;
  return hip_ok(c, hipStreamSynchronize(c.stream), "set_pour_source") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_delete_particles(mpf_ctx* ctx, int64_t* deleted) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  int64_t d = 0, fd = 0;
  int st = run_delete_particles(c, &d, &fd);
  if (deleted) *deleted = d;
  if (st != MPF_OK) return (mpf_status)st;
  return check_launch(c);
}

mpf_status mpf_pour(mpf_ctx* ctx, int32_t frame, int64_t* added) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  int64_t a = 0;
  int st = run_pour(c, frame, &a);
  if (added) *added = a;
  if (st != MPF_OK) return (mpf_status)st;
  mpf_status ls = check_launch(c);
  if (ls != MPF_OK) return ls;
  return hip_ok(c, hipStreamSynchronize(c.stream), "pour") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_hash_sort(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_hash_sort(c);
  return check_launch(c);
}

mpf_status mpf_apply_forces(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_apply_forces(c);
  return check_launch(c);
}

#define REQUIRE_SORTED(c)                                                                  \
  if (!(c).sorted) return fail((c), MPF_INVALID_ARG, "particles not sorted: call mpf_hash_sort first");

mpf_status mpf_mark_solo_air(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_mark_solo_air(c);
  return check_launch(c);
}

mpf_status mpf_surface_tension(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_surface_tension(c);
  return check_launch(c);
}

mpf_status mpf_drag_force(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_drag(c);
  return check_launch(c);
}

mpf_status mpf_update_solo_air(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_solo_air_sph(c);
  return check_launch(c);
}

mpf_status mpf_correct_positions(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_correct_pos(c);
  return check_launch(c);
}

mpf_status mpf_correct_bubble_positions(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_correct_bubble_pos(c);
  return check_launch(c);
}

mpf_status mpf_solid_motion(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  int st = run_solid_motion(c);
  if (st != MPF_OK) return (mpf_status)st;
  return check_launch(c);
}

mpf_status mpf_levelset_update(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_levelset(c);
  return check_launch(c);
}

mpf_status mpf_p2g(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_p2g(c);
  return check_launch(c);
}

mpf_status mpf_extrapolate_bc(mpf_ctx* ctx, int32_t recompute_phi) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_extrapolate_bc(c, recompute_phi != 0);
  return check_launch(c);
}

mpf_status mpf_project(mpf_ctx* ctx, mpf_solve_info* info) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  c.note.clear();
  run_divergence(c);
  int st = run_pressure_solve(c, info);
  // any failure other than stopping at the cap leaves the pressure unusable:
  // return its status before the gradient step touches the face velocities
  if (st != MPF_OK && st != MPF_NO_CONVERGENCE) return (mpf_status)st;
  run_subgrad(c);
  mpf_status ls = check_launch(c);
  if (ls != MPF_OK) return ls;
  if (st == MPF_NO_CONVERGENCE) c.err = c.note.empty() ? "pressure solve stopped at its iteration cap" : c.note;
  return (mpf_status)st;
}

mpf_status mpf_g2p(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  run_g2p(c);
  return check_launch(c);
}

mpf_status mpf_advect(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (c.cfg.integrator != MPF_ADVECT_RK2) REQUIRE_SORTED(c);
  run_advect(c);
  return check_launch(c);
}

#define ADJ_PRELUDE(ctx)                                                                       \
  if (!ctx) return MPF_INVALID_ARG;                                                           \
  Ctx& c = ctx->c;                                                                            \
  c.err.clear();                                                                              \
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");

mpf_status mpf_project_adjoint(mpf_ctx* ctx, const double* const* g_out, double* const* g_in, int32_t where,
                               mpf_solve_info* info) {
  ADJ_PRELUDE(ctx)
  if (!g_in) return fail(c, MPF_INVALID_ARG, "g_in is NULL");
  int st = run_project_adjoint(c, g_out, g_in, where, info);
  if (st == MPF_NO_CONVERGENCE) c.err = "adjoint solve stopped at its iteration cap";
  return (mpf_status)st;
}

mpf_status mpf_extrapolate_adjoint(mpf_ctx* ctx, const double* const* g_out, double* const* g_in, int32_t where) {
  ADJ_PRELUDE(ctx)
  if (!g_in) return fail(c, MPF_INVALID_ARG, "g_in is NULL");
  return (mpf_status)run_extrapolate_adjoint(c, g_out, g_in, where);
}

mpf_status mpf_p2g_adjoint(mpf_ctx* ctx, const double* const* g_faces, double* g_vel, double* g_pos,
                           int32_t where) {
  ADJ_PRELUDE(ctx)
  REQUIRE_SORTED(c);
  return (mpf_status)run_p2g_adjoint(c, g_faces, g_vel, g_pos, where);
}

mpf_status mpf_g2p_adjoint(mpf_ctx* ctx, const double* g_vel_out, double* g_vel_in, double* const* g_u,
                           double* const* g_uold, double* g_pos, int32_t where) {
  ADJ_PRELUDE(ctx)
  return (mpf_status)run_g2p_adjoint(c, g_vel_out, g_vel_in, g_u, g_uold, g_pos, where);
}

mpf_status mpf_advect_adjoint(mpf_ctx* ctx, const double* g_pos_out, const double* g_vel_out, double* g_pos_in,
                              double* g_vel_in, int32_t where, int64_t* in_band) {
  ADJ_PRELUDE(ctx)
  REQUIRE_SORTED(c);
  return (mpf_status)run_advect_adjoint(c, g_pos_out, g_vel_out, g_pos_in, g_vel_in, where, in_band);
}

mpf_status mpf_heat(mpf_ctx* ctx, mpf_solve_info* info) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  run_heat_grid_pre(c);
  int st = MPF_OK;
  if (c.cfg.heat_solver == MPF_HEAT_CG) {
    st = run_heat_cg(c, info);
    if (st != MPF_OK && st != MPF_NO_CONVERGENCE) return (mpf_status)st;
  } else {
    run_heat_jacobi(c, c.cfg.heat_jacobi_iters);
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->iterations = c.cfg.heat_jacobi_iters;
      info->converged = 1;
    }
    st = MPF_OK;
  }
  run_heat_grid_post(c);
  run_heat_g2p(c);
  mpf_status ls = check_launch(c);
  if (ls != MPF_OK) return ls;
  return MPF_OK;  // the reference never fails on the heat cap (spray.cu:1803-1810)
}

mpf_status mpf_phase_change(mpf_ctx* ctx, int32_t frame, mpf_phase_info* info) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  REQUIRE_SORTED(c);
  int st = run_phase_change(c, frame, info);
  if (st != MPF_OK) return (mpf_status)st;
  return check_launch(c);
}

mpf_status mpf_mc_hash_sort(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (!c.mc_cstart)
    return fail(c, MPF_INVALID_ARG, "mpf_mc_hash_sort needs mc_enabled and a particle mc_source at mpf_create");
  run_hash_sort_mc(c);
  return check_launch(c);
}

mpf_status mpf_marching_cubes(mpf_ctx* ctx, mpf_mc_info* info) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (!c.cfg.mc_enabled) return fail(c, MPF_INVALID_ARG, "marching cubes disabled in the configuration");
  if (c.cfg.mc_source != MPF_MC_LEVELSET && !c.mc_sorted)
    return fail(c, MPF_INVALID_ARG, "particles not in 2N order: call mpf_mc_hash_sort first");
  int st = run_marching_cubes(c, info);
  if (st != MPF_OK) return (mpf_status)st;
  return check_launch(c);
}

mpf_status mpf_get_mesh(mpf_ctx* ctx, float* verts, uint32_t* tris, int32_t where) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (verts && c.mc_nv) hipMemcpyAsync(verts, c.mc_verts, (size_t)c.mc_nv * 12, kind, c.stream);
  if (tris && c.mc_nt) hipMemcpyAsync(tris, c.mc_tris, (size_t)c.mc_nt * 12, kind, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_mesh") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_write_mesh(mpf_ctx* ctx, const char* path, int32_t format, const char* object_name) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (!path) return fail(c, MPF_INVALID_ARG, "null path");
  if (format < MPF_MESH_POVRAY || format > MPF_MESH_BOBJ_GZ) return fail(c, MPF_INVALID_ARG, "unknown mesh format");
  if (!c.mc_verts) return fail(c, MPF_INVALID_ARG, "marching cubes disabled in the configuration");
  return (mpf_status)run_write_mesh(c, path, format, object_name);
}

mpf_status mpf_get_mesh_normals(mpf_ctx* ctx, float* normals, int32_t where) {
  if (!ctx || !normals) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (!c.mc_normals) return fail(c, MPF_INVALID_ARG, "marching cubes disabled in the configuration");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (c.mc_nv) hipMemcpyAsync(normals, c.mc_normals, (size_t)c.mc_nv * 12, kind, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_mesh_normals") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_get_mc_cases(mpf_ctx* ctx, uint8_t* cases, int32_t where) {
  if (!ctx || !cases) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (!c.mc_case) return fail(c, MPF_INVALID_ARG, "marching cubes disabled in the configuration");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (c.mc_nvox) hipMemcpyAsync(cases, c.mc_case, (size_t)c.mc_nvox, kind, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_mc_cases") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_get_mc_field(mpf_ctx* ctx, float* field, int32_t where) {
  if (!ctx || !field) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (!c.mc_field) return fail(c, MPF_INVALID_ARG, "no particle marching cubes source in the configuration");
  if (!c.mc_field_valid) return fail(c, MPF_INVALID_ARG, "no particle-source mesh built yet");
  int64_t n = 0, v = 0;
  mc_lattice(c, n, v);
  hipMemcpyAsync(field, c.mc_field, (size_t)n * 4, where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                 c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_mc_field") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_get_mc_cells(mpf_ctx* ctx, uint32_t* cell_start, uint32_t* cell_end, int32_t where) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  if (where != MPF_HOST && where != MPF_DEVICE) return fail(c, MPF_INVALID_ARG, "bad memory location");
  if (!c.mc_cstart) return fail(c, MPF_INVALID_ARG, "no particle marching cubes source in the configuration");
  hipMemcpyKind kind = where == MPF_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (cell_start) hipMemcpyAsync(cell_start, c.mc_cstart, (size_t)c.g2.G * 4, kind, c.stream);
  if (cell_end) hipMemcpyAsync(cell_end, c.mc_cend, (size_t)c.g2.G * 4, kind, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "get_mc_cells") ? MPF_OK : MPF_HIP_ERROR;
}

mpf_status mpf_step(mpf_ctx* ctx, int32_t frame, mpf_step_info* info) {
  if (!ctx) return MPF_INVALID_ARG;
  Ctx& c = ctx->c;
  c.err.clear();
  mpf_step_info loc;
  std::memset(&loc, 0, sizeof(loc));
  mpf_status st;
  if ((st = mpf_update_seed_cells(ctx, frame)) != MPF_OK) return st;
  if ((st = mpf_pour(ctx, frame, &loc.poured)) != MPF_OK) return st;
  if ((st = mpf_hash_sort(ctx)) != MPF_OK) return st;
  if ((st = mpf_apply_forces(ctx)) != MPF_OK) return st;
  if ((st = mpf_levelset_update(ctx)) != MPF_OK) return st;
  const mpf_config& k = c.cfg;
  if (k.solo_air_enabled && (st = mpf_mark_solo_air(ctx)) != MPF_OK) return st;
  if (k.surface_tension_sigma != 0.0f && (st = mpf_surface_tension(ctx)) != MPF_OK) return st;
  if ((st = mpf_p2g(ctx)) != MPF_OK) return st;
  if (k.drag_enabled && (st = mpf_drag_force(ctx)) != MPF_OK) return st;
  if ((st = mpf_extrapolate_bc(ctx, 1)) != MPF_OK) return st;
  mpf_status pst = mpf_project(ctx, &loc.project);
  if (pst != MPF_OK && pst != MPF_NO_CONVERGENCE) return pst;
  const std::string perr = c.err;
  if ((st = mpf_extrapolate_bc(ctx, 0)) != MPF_OK) return st;
  if ((st = mpf_g2p(ctx)) != MPF_OK) return st;
  // bubblesim runs solidmotion before updateSoloAirParticle; the SPH reads
  // only air particles, which the solid motion does not touch, so running it
  // first gives the same result while the cell lists still hold
  if (k.solo_air_enabled && (st = mpf_update_solo_air(ctx)) != MPF_OK) return st;
  if ((st = mpf_solid_motion(ctx)) != MPF_OK) return st;
  // bubblesim sorts before CollisionSolid (spray.cpp:1122); only a moving
  // solid has moved particles since the step's first sort
  if (!c.sorted && (st = mpf_hash_sort(ctx)) != MPF_OK) return st;
  if ((st = mpf_advect(ctx)) != MPF_OK) return st;
  if ((st = mpf_hash_sort(ctx)) != MPF_OK) return st;
  if (k.correct_position) {
    // spray.cpp:1130-1142: spring correction, level set of the corrected
    // particles, bubble-side correction, and the sort for the heat stage
    if ((st = mpf_correct_positions(ctx)) != MPF_OK) return st;
    if ((st = mpf_hash_sort(ctx)) != MPF_OK) return st;
    if ((st = mpf_levelset_update(ctx)) != MPF_OK) return st;
    if ((st = mpf_correct_bubble_positions(ctx)) != MPF_OK) return st;
    if ((st = mpf_hash_sort(ctx)) != MPF_OK) return st;
  } else {
    if ((st = mpf_levelset_update(ctx)) != MPF_OK) return st;
  }
  if (c.cfg.heat_enabled && (st = mpf_heat(ctx, &loc.heat)) != MPF_OK) return st;
  {
    // nucleation, deletion, melting and freezing in the reference's order
    c.err.clear();
    REQUIRE_SORTED(c);
    int64_t deleted = 0;
    int pst2 = run_phase_change_with_delete(c, frame, &loc.phase, &deleted);
    if (pst2 != MPF_OK) return (mpf_status)pst2;
    if ((st = check_launch(c)) != MPF_OK) return st;
    loc.phase.deleted = (int32_t)deleted;
  }
  if (c.cfg.mc_enabled && frame % c.cfg.mc_every == 0) {
    // bubblesim's rendering stage (spray.cpp:1192-1201): preMC, then the mesh
    if (c.cfg.mc_source != MPF_MC_LEVELSET && (st = mpf_mc_hash_sort(ctx)) != MPF_OK) return st;
    if ((st = mpf_marching_cubes(ctx, &loc.mc)) != MPF_OK) return st;
  }
  loc.particles = c.P;
  if (info) *info = loc;
  if (pst == MPF_NO_CONVERGENCE) c.err = perr;
  return pst;
}

mpf_status mpf_set_kernel_timing(mpf_ctx* ctx, int32_t on) {
  if (!ctx) return MPF_INVALID_ARG;
  ctx->c.flush_timing();
  ctx->c.timing = on != 0;
  return MPF_OK;
}

int32_t mpf_kernel_times(mpf_ctx* ctx, char* names, int64_t* counts, double* total_ms, int32_t cap) {
  if (!ctx) return -1;
  Ctx& c = ctx->c;
  c.flush_timing();
  int32_t n = (int32_t)c.stats.size();
  for (int32_t i = 0; i < n && i < cap; ++i) {
    if (names) {
      std::memset(names + 48 * i, 0, 48);
      std::strncpy(names + 48 * i, c.stats[i].name.c_str(), 47);
    }
    if (counts) counts[i] = c.stats[i].count;
    if (total_ms) total_ms[i] = c.stats[i].ms;
  }
  return n;
}

mpf_status mpf_reset_kernel_times(mpf_ctx* ctx) {
  if (!ctx) return MPF_INVALID_ARG;
  ctx->c.flush_timing();
  ctx->c.stats.clear();
  return MPF_OK;
}

}  // extern "C"
