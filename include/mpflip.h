/*
 * mpflip.h — C ABI of the MI355X-native FLIP multiphase step.
 *
 * The reference's "plugin API" for this path is the set of argument-less host
 * step methods of `class cspray` (Multiphase/Multiphase/spray.h:18-389), which
 * operate on member state (device buffers owned by the object) and are driven
 * by `cspray::bubblesim()` (spray.cpp:979-1233).  This header keeps that shape:
 * an opaque context owns every device buffer and one HIP stream, and each
 * entry point below replaces one of those methods.  Everything crossing the
 * boundary is a plain pointer, a size or a POD struct; no C++ or torch types.
 *
 * Layouts accepted (the reference's, spray.h:165-169, utility.h:49-88):
 *   particles  float pos[3*P], vel[3*P] (AoS float3), int8 flag[P],
 *              float mass[P], T[P], LHeat[P], solubility[P], gascontain[P]
 *   cells      x-major, index i*NY*NZ + j*NZ + k (z contiguous)
 *   faces      ux[(NX+1)*NY*NZ], uy[NX*(NY+1)*NZ], uz[NX*NY*(NZ+1)], same order
 *   cell lists uint32 cell_start[G], cell_end[G]; empty cell = MPF_CELL_UNDEF
 *
 * Every function returns mpf_status.  A context is not thread-safe; distinct
 * contexts are independent.  Pointers passed in are borrowed for the call.
 */
#ifndef MPFLIP_H
#define MPFLIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPF_ABI_VERSION 2
#define MPF_CELL_UNDEF 0xffffffffu

/* Particle / cell type codes: utility.h:13-19. */
enum {
  MPF_TYPEFLUID = 0,
  MPF_TYPEAIR = 1,
  MPF_TYPEBOUNDARY = 2,
  MPF_TYPEVACUUM = 3,
  MPF_TYPEAIRSOLO = 4,
  MPF_TYPESOLID = 5,
  MPF_TYPECNT = 6
};

typedef enum {
  MPF_OK = 0,
  MPF_INVALID_ARG = 1,
  MPF_NONFINITE = 2,
  MPF_NO_CONVERGENCE = 3,
  MPF_CAPACITY = 4,
  MPF_HIP_ERROR = 5,
  MPF_IO = 6
} mpf_status;

/* Where a caller buffer lives. */
enum { MPF_HOST = 0, MPF_DEVICE = 1 };

/* Arithmetic of the pressure projection (divergence, PCG, gradient). */
enum {
  MPF_FP32 = 0,  /* FP32 vectors, FP32 block sums of the dot products        */
  MPF_FP64 = 1,  /* FP64 divergence, vectors, dots and pressure               */
  MPF_MIXED = 2  /* FP32 vectors and SpMV, FP64 dots, FP64 pressure updated
                    and true residual recomputed in FP64 every refine_every
                    iterations (reliable update)                              */
};

/* Pressure solver. */
enum {
  MPF_SOLVER_REFERENCE_CG = 0, /* spray.cu:1613-1687: no preconditioner, stop
                                  when r.r < 1e-2*G (absolute), cap ref_max_iters */
  MPF_SOLVER_JACOBI_PCG = 1    /* Jacobi PCG, stop when |r|/|b| < tol        */
};

/* Pressure operator. */
enum {
  MPF_OPERATOR_UNIT = 0,   /* computeAx_bubble, spray_k.cu:3423-3447          */
  MPF_OPERATOR_VARDENS = 1 /* 1/rho face weights; equals UNIT at rho_l=rho_a=1 */
};

/* Particle integrator. */
enum {
  MPF_ADVECT_EULER_COLLIDE = 0, /* CollisionWithSolid_k, spray_k.cu:5276-5345 */
  MPF_ADVECT_RK2 = 1,           /* advectparticle_RK2,   spray_k.cu:409-459    */
  MPF_ADVECT_FREEZING = 2       /* CollisionWithSolid_Freezing, spray_k.cu:5348-5446
                                   (the freezing and melting-pour scenes)    */
};

/* Field that marching cubes meshes (mpf_config.mc_source; runMC_smooth,
   spray.cu:802-826). */
enum {
  MPF_MC_LEVELSET = 0,   /* lsfluid at cell centres                           */
  MPF_MC_LIQUID_GAS = 1, /* runMC_fluid: genWaterDensfield_liquidAndGas,
                            spray_k.cu:1595-1686, FLUID/AIR/AIRSOLO particles  */
  MPF_MC_GAS = 2,        /* runMC_gas: genWaterDensfield_Gas,
                            spray_k.cu:1528-1593, AIR (+AIRSOLO) particles     */
  MPF_MC_SOLID = 3       /* runMC_solid: genWaterDensfield_GY,
                            spray_k.cu:1688-1772, SOLID particles              */
};

/* Heat diffusion solver. */
enum {
  MPF_HEAT_CG = 0,    /* solver_cg_heat, spray.cu:1780-1852                  */
  MPF_HEAT_JACOBI = 1 /* fixed number of Jacobi sweeps on the same operator  */
};

typedef struct mpf_config {
  int32_t nx, ny, nz;
  float dx;            /* cell size (the reference fixes 1/64, spray.cpp:78)  */
  float dt;
  float gravity[3];
  float gmin[3];       /* domain origin; gmax = gmin + dx*(nx,ny,nz)          */
  float samplespace;   /* particle spacing, dx/2 (spray.cpp:79)                */
  float m0, airm0;     /* particle masses (spray.cpp:581-582)                  */
  int64_t particle_capacity;

  /* forces (spray.cu:376-393) */
  int32_t buoyancy_kick;     /* addbuoyancyforce_vel (reference mode)         */
  float bubble_max_vel, buoyancy_rate_air, buoyancy_rate_solo;

  /* projection */
  int32_t precision;         /* MPF_FP32 / MPF_FP64 / MPF_MIXED               */
  int32_t solver;            /* MPF_SOLVER_*                                  */
  int32_t op;                /* MPF_OPERATOR_*                                */
  float rho_liquid, rho_air; /* densities for MPF_OPERATOR_VARDENS            */
  double tol;                /* relative residual for MPF_SOLVER_JACOBI_PCG   */
  int32_t max_iters;         /* PCG iteration cap                             */
  int32_t ref_max_iters;     /* reference CG cap (spray.cu:1642, 800)         */
  int32_t refine_every;      /* MPF_MIXED: true-residual interval (20)        */
  int32_t check_every;       /* host convergence poll interval (iterations)   */
  int32_t solid_constraint;  /* cells dominated by SOLID particles become
                                velocity-constrained (Neumann) in the solve   */
  int32_t flags_from_counts; /* markgrid_bubble (MELTING/FREEZING scenes)     */

  /* FLIP / advection */
  float flip_alpha;          /* 1 = pure FLIP (bubblesim), 0.95 = watersim    */
  int32_t integrator;        /* MPF_ADVECT_*                                  */
  int32_t fix_solids;        /* solidmotion_fixed (spray.cu:2685-2689)        */
  float bounce_vel, bounce_pos;

  /* heat (spray.cpp:35-38, 62-65; spray.cu:3076-3090) */
  int32_t heat_enabled;
  int32_t heat_solver;       /* MPF_HEAT_*                                    */
  int32_t heat_jacobi_iters; /* 10                                            */
  int32_t heat_max_iters;    /* CG cap (3800)                                 */
  float heat_alpha[MPF_TYPECNT]; /* diffusivity per cell type                 */
  float default_temperature; /* T of cells without particles                  */
  int32_t heat_add_bottom;   /* addHeatAtBottom                               */
  float heat_increase_bottom, heat_bottom_clamp;
  int32_t heat_extend_boundary; /* setBoundaryHeat                            */
  int32_t heat_vacuum_neumann;  /* boiling scenes: VACUUM neighbours Neumann  */
  int32_t heat_melt_and_boil;   /* mapHeatg2p_MeltAndBoil                     */
  float alpha_temp_trans;

  /* phase change */
  int32_t melt_enabled, freeze_enabled, gas_enabled;
  int32_t gas_every;           /* GenerateGasParticle every n frames (2)      */
  float melting_point, boiling_point, liquid_heat_th;
  float temperature0, dissolve_gas_rate, vapor_gen_rate;

  /* marching cubes on the liquid level set */
  int32_t mc_enabled;
  int64_t mc_max_verts, mc_max_tris;

  uint64_t seed;

  /* ---- appended in ABI version 2 ---- */
  int32_t adjoint_enabled;   /* allocate the adjoint entry points' scratch at
                                mpf_create (the adjoints refuse otherwise)    */
  /* particle deletion (deleteAirFluidParticle, spray.cu:2194-2248) */
  int32_t delete_enabled;
  int32_t delete_vacuum_sweeps; /* sweepVacuum passes (20)                    */
  /* pouring (pouring / pouringgas, spray.cu:3012-3043) */
  int32_t pour_every;        /* frames between pours (4; 6 for gas)           */
  float pour_pos_rand;       /* posrandparam (0.1, spray.cpp:55), x samplespace */
  float pour_vel_rand;       /* velrandparam (0.0, spray.cpp:55)              */
  float pour_temperature;    /* defaultLiquidT (293.15, spray.cpp:54)         */
  /* nucleation seed refresh (updateSeedCell, spray.cu:2933-2950) */
  int32_t seed_update_every; /* frames between refreshes, 0 = never (15)      */
  /* marching cubes source: MPF_MC_LEVELSET, or a particle field on the
     (2nx+1)(2ny+1)(2nz+1) node lattice of spacing dx/2 (preMC + runMC_*,
     spray.cu:759-826), whose buffers are allocated at mpf_create */
  int32_t mc_source;
  int32_t mc_gas_airsolo;    /* MPF_MC_GAS counts AIRSOLO particles (1; the
                                reference's SCENE_INTERACTION uses 0)         */
  int32_t mc_every;          /* mpf_step meshes frames that are multiples of
                                this (1; the reference's outputframeDelta is 2) */
  int32_t mc_smooth_iters;   /* smoothMesh iterations after meshing (5;
                                m_bSmoothMC / smoothIterTimes, spray.cpp:62, 610) */
  /* remaining bubblesim stages (spray.cpp:1059-1150), reference defaults */
  int32_t solo_air_enabled;  /* markSoloAirParticle + updateSoloAirParticle (1) */
  float solo_tol_in, solo_tol_out; /* verifySoloAirParticle tol1, tol2 in cells
                                (-1.45, -0.5; MELTANDBOIL 0.05, -0.8)         */
  float sph_viscosity;       /* viscosiySPH (1.8)                             */
  float surface_tension_sigma; /* surfacetensionsigma (1.0; 0 skips the stage) */
  int32_t drag_enabled;      /* enforceDragForce (1)                          */
  float drag_solo, drag_grid; /* dragParamSolo, dragParamGrid (10, 0.05)      */
  int32_t drag_particle_update; /* the drag also accelerates the solo air
                                particle (SCENE_INTERACTION; 0)               */
  int32_t correct_position;  /* m_bCorrectPosition (1)                        */
  float correction_spring, correction_radius; /* (500000, 0.5 cells)          */
  int32_t correct_bubble_air_ls; /* correctbubblepos_air (the MELTANDBOIL,
                                INTERACTION and ALL scenes; 0)                */
  /* rigid solid motion (solidmotion, spray.cu:2691-2774; fix_solids = 0) */
  float solid_bounce;        /* SolidbounceParam (0.5, spray.cpp:52)          */
  int32_t rigid_lock_y;      /* SCENE_INTERACTION: no translation along y (0) */
  /* MPF_MIXED single-kernel iteration without slow rows: 1 forms q_{k-1} =
     A p_{k-1} again from p_{k-1} instead of storing and reading q (24 bytes
     per unknown and iteration instead of 32, iterates equal to rounding;
     on MI355X it is bound by its arithmetic and runs within 2% of the
     stored-q kernel); 0 stores q (0) */
  int32_t pcg_recompute_q;
} mpf_config;

/* Opaque context. */
typedef struct mpf_ctx mpf_ctx;

/* Grid / particle fields addressable by mpf_get_field / mpf_set_field. */
typedef enum {
  /* cell fields (G = nx*ny*nz) */
  MPF_F_MARK = 0,     /* int8   */
  MPF_F_LSFLUID, MPF_F_LSAIR, MPF_F_LSMERGE,
  MPF_F_PHIGRAD_X, MPF_F_PHIGRAD_Y, MPF_F_PHIGRAD_Z,
  MPF_F_PHIFLUID, MPF_F_PHIAIR, MPF_F_PHISOLID,
  MPF_F_DIV,          /* float, or double under MPF_FP64                      */
  MPF_F_PRESSURE,     /* float under MPF_FP32, double under MPF_FP64/MIXED    */
  MPF_F_TEMPERATURE, MPF_F_TEMPERATURE_CHANGE,
  MPF_F_CELL_START, MPF_F_CELL_END, /* uint32 */
  /* face fields */
  MPF_F_WATER_UX, MPF_F_WATER_UY, MPF_F_WATER_UZ,
  MPF_F_AIR_UX, MPF_F_AIR_UY, MPF_F_AIR_UZ,
  MPF_F_WATER_UX_OLD, MPF_F_WATER_UY_OLD, MPF_F_WATER_UZ_OLD,
  MPF_F_AIR_UX_OLD, MPF_F_AIR_UY_OLD, MPF_F_AIR_UZ_OLD,
  MPF_F_SOLID_UX, MPF_F_SOLID_UY, MPF_F_SOLID_UZ,
  /* particle fields (P = current count) */
  MPF_F_GRID_HASH, MPF_F_GRID_INDEX, /* uint32, from the last mpf_hash_sort  */
  MPF_F_SURFACE_TENSION,             /* cell float, compsurfacetension_k      */
  MPF_F_COUNT
} mpf_field;

typedef struct mpf_solve_info {
  int32_t iterations;      /* solver iterations                               */
  int32_t refinements;     /* FP64 true-residual recomputations (MIXED)       */
  int32_t converged;
  int32_t pad_;
  double rel_residual;     /* |b - A x| / |b|, true residual, FP64            */
  double abs_rr;           /* recurrence r.r at exit                          */
} mpf_solve_info;

typedef struct mpf_phase_info {
  int32_t melted, frozen, gas_added;
  int32_t deleted;         /* mpf_step with delete_enabled: particles deleted */
  int64_t particles;
} mpf_phase_info;

typedef struct mpf_mc_info {
  int64_t active_voxels, verts, tris;
} mpf_mc_info;

typedef struct mpf_step_info {
  mpf_solve_info project;
  mpf_solve_info heat;
  mpf_phase_info phase;
  mpf_mc_info mc;
  int64_t particles;
  int64_t poured;          /* particles appended by the pour source           */
} mpf_step_info;

/* ---- lifecycle --------------------------------------------------------- */
/* Reference defaults of initparam (spray.cpp:23-79, 573-618) for an nx*ny*nz
   grid.  Replaces the compile-time NX/NY/NZ (spray.h:9-16) and __constant__
   FlipConstant dparam (spray_k.cu:8, utility.h:26-46). */
mpf_status mpf_config_default(mpf_config* cfg, int32_t nx, int32_t ny, int32_t nz);
/* Replaces cspray::init()/initmem_bubble() (spray.cpp:628-670, spray.cu:25-218). */
mpf_status mpf_create(const mpf_config* cfg, mpf_ctx** out);
mpf_status mpf_destroy(mpf_ctx* ctx);
/* Message of the last failing call on ctx (or, when ctx is NULL, of the last
   failing context-free call: mpf_create, mpf_hash_sort_noalloc). */
const char* mpf_last_error(mpf_ctx* ctx);
int32_t mpf_abi_version(void);
/* The context's HIP stream (hipStream_t), for callers that time or order work. */
void* mpf_stream(mpf_ctx* ctx);
mpf_status mpf_synchronize(mpf_ctx* ctx);

/* ---- data movement ----------------------------------------------------- */
/* Replaces initparticle_* H2D copies (spray.cu:395-484). Any pointer may be
   NULL, which zero-fills that attribute (flag NULL = TYPEFLUID). */
mpf_status mpf_set_particles(mpf_ctx* ctx, int64_t n, const float* pos, const float* vel,
                             const int8_t* flag, const float* mass, const float* T,
                             const float* lheat, const float* solubility,
                             const float* gascontain, int32_t where);
/* Any output pointer may be NULL (skipped). */
mpf_status mpf_get_particles(mpf_ctx* ctx, int64_t* n, float* pos, float* vel, int8_t* flag,
                             float* mass, float* T, float* lheat, float* solubility,
                             float* gascontain, int32_t where);
int64_t mpf_particle_count(mpf_ctx* ctx);
/* Element count and element size in bytes of a field. */
mpf_status mpf_field_info(mpf_ctx* ctx, int32_t field, int64_t* count, int32_t* elem_bytes);
mpf_status mpf_get_field(mpf_ctx* ctx, int32_t field, void* dst, int32_t where);
mpf_status mpf_set_field(mpf_ctx* ctx, int32_t field, const void* src, int32_t where);
/* Nucleation seed cells (cell indices), spray.cu:2907-2951. */
mpf_status mpf_set_seed_cells(mpf_ctx* ctx, const int32_t* cells, int32_t n);
/* cspray::updateSeedCell, spray.cu:2933-2950: when frame is a multiple of
   cfg.seed_update_every, four seed cells (a ring over the list) are replaced
   by random cells at k = 2, 3 <= i < nx-3, 3 <= j < ny-3 (counter-hash random
   numbers of (seed, frame, slot)).  mpf_step calls it first. */
mpf_status mpf_update_seed_cells(mpf_ctx* ctx, int32_t frame);
/* Pour source of pouring/pouringgas (spray.cu:3012-3043, positions built by
   CompPouringParam_*, spray.cpp:1415-1545): n positions and velocities
   (float[3n], host memory) and the type of the poured particles (FLUID, or
   AIRSOLO for pouringgas).  n = 0 removes the source. */
mpf_status mpf_set_pour_source(mpf_ctx* ctx, int64_t n, const float* pos, const float* vel, int8_t flag);

/* ---- step entry points (one per reference host step method) ------------ */
/* cspray::hashAndSortParticles, spray.cu:614-637: calcHashD + stable sort by
   (key, index) + reorderDataAndFindCellStartD. */
mpf_status mpf_hash_sort(mpf_ctx* ctx);
/* cspray::addexternalforce, spray.cu:376-393. */
mpf_status mpf_apply_forces(mpf_ctx* ctx);
/* cspray::computeLevelset + sweepLSAndMardGrid, spray.cu:1465-1471, 2011-2065
   (deterministic two-buffer sweeps); then markgrid_bubble (spray.cu:522-529)
   when cfg.flags_from_counts, and the solid marking when cfg.solid_constraint. */
mpf_status mpf_levelset_update(mpf_ctx* ctx);
/* cspray::mapvelp2g_bubble, spray.cu:1562-1588 (both phases + FLIP snapshot). */
mpf_status mpf_p2g(mpf_ctx* ctx);
/* sweepPhi/sweepU/setWaterBoundaryU for both phases, spray.cpp:1088-1094
   (recompute_phi = 1) or spray.cpp:1101-1104 (recompute_phi = 0). */
mpf_status mpf_extrapolate_bc(mpf_ctx* ctx, int32_t recompute_phi);
/* cspray::project_CG_bubble, spray.cu:1590-1611.  MPF_NO_CONVERGENCE when the
   solver stops at its cap without meeting its criterion (results are kept). */
mpf_status mpf_project(mpf_ctx* ctx, mpf_solve_info* info);
/* cspray::mapvelg2p_bubble, spray.cu:1689-1705, with blend cfg.flip_alpha. */
mpf_status mpf_g2p(mpf_ctx* ctx);
/* cspray::solidmotion_fixed + CollisionSolid, spray.cu:2685-2689, 2959-2983
   (MPF_ADVECT_EULER_COLLIDE) or wateradvect, spray.cu:293-305 (MPF_ADVECT_RK2). */
mpf_status mpf_advect(mpf_ctx* ctx);
/* cspray::updateTemperature, spray.cu:1889-1947 (incl. updateLatentHeat). */
mpf_status mpf_heat(mpf_ctx* ctx, mpf_solve_info* info);
/* GenerateGasParticle, MeltSolid, Freezing: spray.cu:2653-2683, 2816-2875. */
mpf_status mpf_phase_change(mpf_ctx* ctx, int32_t frame, mpf_phase_info* info);
/* cspray::deleteAirFluidParticle, spray.cu:2194-2248, on the cell flags and
   air level set of the last mpf_levelset_update: cfg.delete_vacuum_sweeps
   passes turn AIR cells with a VACUUM corner-diagonal neighbour into VACUUM
   (two-buffer passes), then FLUID particles with lsair < -h and AIR/AIRSOLO
   particles with a VACUUM cell in their 3x3x3 neighbourhood (or mass <= 1e-6)
   are removed, order preserved (stable compaction).  Cell lists that indexed
   every particle are remapped and stay valid.  *deleted (may be NULL): the
   number removed. */
mpf_status mpf_delete_particles(mpf_ctx* ctx, int64_t* deleted);
/* cspray::pouring / pouringgas: when frame is a multiple of cfg.pour_every
   and the particles fit, append the pour source's particles (random offsets
   cfg.pour_pos_rand * samplespace and cfg.pour_vel_rand per component).
   *added (may be NULL): particles appended. */
mpf_status mpf_pour(mpf_ctx* ctx, int32_t frame, int64_t* added);
/* cspray::preMC, spray.cu:759-765 (hashAndSortParticles_MC, spray.cu:639-662;
   calcHashD_MC, spray_k.cu:1030-1048): the hash, stable sort and particle
   reorder of mpf_hash_sort on the 2nx x 2ny x 2nz grid of cell size
   (dx / 2nx) * nx.  Particles come back in that order (MPF_F_GRID_HASH and
   MPF_F_GRID_INDEX hold its keys and permutation); the cell lists of
   mpf_hash_sort are stale until it runs again.  Needs a particle mc_source. */
mpf_status mpf_mc_hash_sort(mpf_ctx* ctx);
/* runMC_smooth (spray.cu:802-979) on cfg.mc_source: for MPF_MC_LEVELSET the
   lattice is the cell centres of lsfluid; for a particle source the node
   field of genWaterDensfield_* is built first from the lists of the last
   mpf_mc_hash_sort, which must be current (MPF_INVALID_ARG otherwise).
   Voxels are classified (classifyVoxel, spray_k.cu:1853-1888), active edges
   numbered in the order of GetEdgeID (spray_k.cu:1925-1957), so vertex and
   triangle indices follow the reference's numbering. */
mpf_status mpf_marching_cubes(mpf_ctx* ctx, mpf_mc_info* info);
/* Copy the last mesh out: verts float[3*nv], tri indices uint32[3*nt].  The
   mesh is smoothed cfg.mc_smooth_iters times (smoothMesh, spray.cu:1271-1288)
   and given vertex normals (calNormals, spray.cu:981-997), each vertex
   summing its triangles in the order of the triangle list. */
mpf_status mpf_get_mesh(mpf_ctx* ctx, float* verts, uint32_t* tris, int32_t where);
mpf_status mpf_get_mesh_normals(mpf_ctx* ctx, float* normals, int32_t where);

/* Mesh files of the last mesh (host I/O after the mesh is copied out). */
enum {
  MPF_MESH_POVRAY = 0,  /* outputPovRaywater, spray.cu:2250-2310: a mesh2 object */
  MPF_MESH_OBJ = 1,     /* outputOBJwater, spray.cu:2312-2372                    */
  MPF_MESH_BOBJ_GZ = 2  /* outputBOBJwater, spray.cu:2373-2435 (gzip level 1)    */
};
/* Writes the mesh (vertices, normals, faces) to path in the given format;
   object_name names the POV-Ray material (may be NULL: "water").  Non-finite
   vertex or normal components are written as 0 (verifyfloat3).  MPF_IO when
   the file cannot be written. */
mpf_status mpf_write_mesh(mpf_ctx* ctx, const char* path, int32_t format, const char* object_name);
/* readdata_solid, LoadSolidData.cpp:12-59: an .asc point cloud (first line
   the point count, then one line per point: an index and three coordinates).
   Up to cap points are stored in xyz (double[3*cap]); *n gets the count the
   file declares.  Context-free: mpf_last_error(NULL) explains a failure. */
mpf_status mpf_read_point_cloud(const char* path, int64_t cap, double* xyz, int64_t* n);
/* Per-voxel case index of the last mesh (uint8): (nx-1)(ny-1)(nz-1) voxels
   over cell centres, or 2nx*2ny*2nz for a particle source. */
mpf_status mpf_get_mc_cases(mpf_ctx* ctx, uint8_t* cases, int32_t where);
/* Node field of the last particle-source mesh, float[(2nx+1)(2ny+1)(2nz+1)]
   (node (i,j,k) at index (i*(2ny+1) + j)*(2nz+1) + k). */
mpf_status mpf_get_mc_field(mpf_ctx* ctx, float* field, int32_t where);
/* Cell lists of the last mpf_mc_hash_sort, uint32[2nx*2ny*2nz] each (either
   pointer may be NULL). */
mpf_status mpf_get_mc_cells(mpf_ctx* ctx, uint32_t* cell_start, uint32_t* cell_end, int32_t where);

/* ---- the remaining bubblesim stages ------------------------------------ */
/* cspray::markSoloAirParticle, spray.cu:2614-2618 (verifySoloAirParticle,
   spray_k.cu:4223-4261): AIR particles with few air neighbours, deep in
   the liquid or near the x/y walls become AIRSOLO (velocity from the air
   grid); AIRSOLO particles with many air neighbours or near the air become
   AIR.  Needs the cell lists and mpf_levelset_update's LSMERGE. */
mpf_status mpf_mark_solo_air(mpf_ctx* ctx);
/* cspray::computesurfacetension, spray.cu:2001-2009: the curvature force
   field sigma * div(n) (compsurfacetension_k, spray_k.cu:3543-3637) and its
   velocity change on AIR particles within one cell of the interface
   (enforcesurfacetension_p, 3639-3680). */
mpf_status mpf_surface_tension(mpf_ctx* ctx);
/* cspray::enforceDragForce, spray.cu:2953-2957 (calDragForce, spray_k.cu:
   5127-5165): each AIRSOLO particle's drag on the liquid faces of its cell,
   added in particle order (the reference adds with atomics).  After mpf_p2g. */
mpf_status mpf_drag_force(mpf_ctx* ctx);
/* cspray::updateSoloAirParticle, spray.cu:2620-2651: SPH density, pressure
   and viscosity of AIR/AIRSOLO particles (calcDensPress_Air,
   enforceForceSoloAirP, spray_k.cu:4284-4373); neighbour velocities are
   the pass's input (the reference reads them while they are rewritten). */
mpf_status mpf_update_solo_air(mpf_ctx* ctx);
/* cspray::solidmotion_fixed / solidmotion, spray.cu:2685-2774: with
   cfg.fix_solids the SOLID particles' velocities are zeroed; otherwise the
   solid moves rigidly: wall push (solidCollisionWithBound, spray_k.cu:
   5478-5508), translation = mean velocity, centre = mean position, rotation
   = the reference's estimate sum(r x v) / mean|r|^2 / n (compute_cI_k), each
   particle's velocity averaged with the rigid one and its position rotated
   and translated (computeVelSolid_k, computePosSolid_k).  Sums are taken in
   double in a fixed order (the reference adds float block sums).  Moves
   particles: a sort must follow. */
mpf_status mpf_solid_motion(mpf_ctx* ctx);
/* cspray::correctpos, spray.cu:328-339 (correctparticlepos, spray_k.cu:
   1357-1411): spring separation of FLUID/AIR particles, clamped to the
   domain.  Leaves the particles unsorted. */
mpf_status mpf_correct_positions(mpf_ctx* ctx);
/* cspray::correctpos_bubble, spray.cu:2067-2098 (correctbubblepos,
   spray_k.cu:3884-3931, or correctbubblepos_air, 3934-3991, when
   cfg.correct_bubble_air_ls): particles on the wrong side of the merged
   level set are pushed across it.  Uses mpf_levelset_update's fields. */
mpf_status mpf_correct_bubble_positions(mpf_ctx* ctx);

/* One full step in bubblesim order (spray.cpp:979-1233), the stages above
   only: seed refresh, pour, sort, forces, level set, solo-air marking,
   surface tension, P2G, drag, extrapolation, projection, extrapolation,
   G2P, solid motion, solo-air SPH, advection, then either (cfg.correct_position) sort,
   position correction, sort, level set, bubble position correction, sort,
   or sort and level set; heat, nucleation, deletion (cfg.delete_enabled),
   melting/freezing, and on frames
   that are multiples of cfg.mc_every marching cubes (after mpf_mc_hash_sort
   for a particle source, as bubblesim calls preMC). */
mpf_status mpf_step(mpf_ctx* ctx, int32_t frame, mpf_step_info* info);

/* ---- adjoints ----------------------------------------------------------
   No counterpart in the reference (it has no differentiation).  Each call is
   the transposed Jacobian of one stage, evaluated at the context's current
   state, which must be the state right after that stage's forward call;
   every discrete decision of that call (flags, level sets, cell lists,
   branch choices) is held fixed.  Cotangents are FP64 arrays in `where`
   memory.  Face sets are 6 pointers: water ux, uy, uz, air ux, uy, uz
   (sizes (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1)); particle arrays are 3P.
   Input cotangent pointers may be NULL (zero); output pointers may be NULL
   (not written).
   Projection (mpf_project, spray.cu:1590-1611): g_in = P^T g_out, one PCG
   solve on the same symmetric operator in cfg.precision. */
mpf_status mpf_project_adjoint(mpf_ctx* ctx, const double* const* g_out, double* const* g_in, int32_t where,
                               mpf_solve_info* info);
/* Extrapolation and wall rule (mpf_extrapolate_bc, spray.cpp:1088-1104). */
mpf_status mpf_extrapolate_adjoint(mpf_ctx* ctx, const double* const* g_out, double* const* g_in, int32_t where);
/* P2G of both phases (mpf_p2g, spray.cu:1562-1588): cotangents of the P2G
   output faces to particle velocities and positions (g_vel, g_pos are
   overwritten). */
mpf_status mpf_p2g_adjoint(mpf_ctx* ctx, const double* const* g_faces, double* g_vel, double* g_pos,
                           int32_t where);
/* G2P / FLIP-PIC blend (mpf_g2p, spray.cu:1689-1705): from the cotangent of
   the new particle velocity to the old velocity, the face velocities u (6),
   the P2G snapshot u_old (6) and positions (all overwritten). */
mpf_status mpf_g2p_adjoint(mpf_ctx* ctx, const double* g_vel_out, double* g_vel_in, double* const* g_u,
                           double* const* g_uold, double* g_pos, int32_t where);
/* Advection (mpf_advect, MPF_ADVECT_EULER_COLLIDE, spray.cu:2959-2983): the
   domain clamp is differentiated; the solid-collision response is not, and
   *in_band (may be NULL) returns how many particles were inside its band. */
mpf_status mpf_advect_adjoint(mpf_ctx* ctx, const double* g_pos_out, const double* g_vel_out, double* g_pos_in,
                              double* g_vel_in, int32_t where, int64_t* in_band);

/* ---- context-free sort without allocation -------------------------------
   No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the cell hash and stable sort of mpf_hash_sort on the
   grid of `cfg` (nx, ny, nz, dx, gmin; nothing else is read), with the
   caller's device buffers.  In: pos float[3n].  Out: grid_hash, grid_index
   uint32[n] (sorted keys, permutation) and cell_start, cell_end uint32[G]
   (MPF_CELL_UNDEF for empty cells), bit-identical to mpf_hash_sort's
   MPF_F_GRID_HASH, MPF_F_GRID_INDEX, MPF_F_CELL_START, MPF_F_CELL_END; the
   particle arrays are not reordered.  Launches on `stream` (a hipStream_t,
   NULL = the null stream) without synchronising; calls no hipMalloc/hipFree.
   Scratch: a fixed device array of 65536 uint32 (256 KiB) per GPU, held
   until the process ends, so G <= 65535 * 4096 cells; larger grids return
   MPF_INVALID_ARG.  Every call on one GPU shares that array: calls that
   overlap on one GPU (different streams or threads) race.  Errors: message
   through mpf_last_error(NULL). */
mpf_status mpf_hash_sort_noalloc(const mpf_config* cfg, int64_t n, const float* pos, uint32_t* grid_hash,
                                 uint32_t* grid_index, uint32_t* cell_start, uint32_t* cell_end, void* stream);

/* ---- introspection ------------------------------------------------------ */
/* Kernel timing with HIP events around every launch on the context stream. */
mpf_status mpf_set_kernel_timing(mpf_ctx* ctx, int32_t on);
/* Up to cap entries: names (48 bytes each, NUL-terminated), launch counts and
   total milliseconds. Returns the number of distinct kernels (may exceed cap). */
int32_t mpf_kernel_times(mpf_ctx* ctx, char* names, int64_t* counts, double* total_ms, int32_t cap);
mpf_status mpf_reset_kernel_times(mpf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MPFLIP_H */
