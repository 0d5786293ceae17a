/*
 * CPU oracle of the FLIP multiphase step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement (scalar loops, optional OpenMP over independent
 * elements, fixed-order reductions) of the reference's kernel bodies in
 * Multiphase/Multiphase/spray_k.cu and host sequencing in spray.cu / spray.cpp.
 * It is used by tests/ as the checker, by __graft_entry__.smoke() as the
 * checker, and by bench.py as the CPU baseline.  The shipped library never
 * links or calls it.
 *
 * The interface mirrors include/mpflip.h on host memory: an opaque state
 * built from the same mpf_config, the same per-stage calls and the same field
 * ids, so a test drives both sides identically.  `use_double` selects the
 * arithmetic: 0 = float (bit-exact partner of the GPU for integer outputs),
 * 1 = double (the FP64 reference of a step).
 *
 * Parity status of each stage and what pins it is listed in tests/ (see the
 * test module docstrings) and in the repository's design notes.
 */
#ifndef MPF_ORACLE_H
#define MPF_ORACLE_H
#include "../include/mpflip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state* orc_create(const mpf_config* cfg, int use_double);
void orc_destroy(orc_state* s);
int orc_set_threads(int n);

/* Particles: float arrays in reference layout (converted to T inside). */
int orc_set_particles(orc_state* s, int64_t n, const float* pos, const float* vel,
                      const int8_t* flag, const float* mass, const float* T,
                      const float* lheat, const float* solubility, const float* gascontain);
/* Outputs are written as double arrays (pos/vel 3n) so no precision is lost. */
int orc_get_particles(orc_state* s, int64_t* n, double* pos, double* vel, int8_t* flag,
                      double* mass, double* T, double* lheat, double* solubility,
                      double* gascontain);
/* Overwrite positions / velocities (3n doubles, current particle order; either
   may be NULL) and nothing else: cell lists, flags and fields are kept.  Used
   for finite differences with the discrete state held fixed. */
int orc_set_kinematics64(orc_state* s, const double* pos, const double* vel);
int64_t orc_particle_count(orc_state* s);
int64_t orc_field_count(orc_state* s, int field);
/* Real fields are exchanged as double; MARK as int8; CELL_*, GRID_* as uint32. */
int orc_get_field(orc_state* s, int field, void* dst);
int orc_set_field(orc_state* s, int field, const void* src);
int orc_set_seed_cells(orc_state* s, const int32_t* cells, int32_t n);
int orc_get_seed_cells(orc_state* s, int32_t* cells); /* returns the count */
int orc_update_seed_cells(orc_state* s, int frame);
int orc_set_pour_source(orc_state* s, int64_t n, const float* pos, const float* vel, int8_t flag);
int64_t orc_pour(orc_state* s, int frame);
/* keep_lists: the cell lists index every particle and are remapped. */
int64_t orc_delete_particles(orc_state* s, int keep_lists);

int orc_hash_sort(orc_state* s);
/* The remaining bubblesim stages (mpf_mark_solo_air ... mpf_correct_bubble_positions). */
int orc_mark_solo_air(orc_state* s);
int orc_surface_tension(orc_state* s);
int orc_drag_force(orc_state* s);
int orc_update_solo_air(orc_state* s);
int orc_correct_positions(orc_state* s);
int orc_correct_bubble_positions(orc_state* s);
int orc_solid_motion(orc_state* s);
/* Vertex normals of the last mesh (double[3*nv]). */
int orc_get_mesh_normals(orc_state* s, double* normals);
/* preMC: the sort on the 2N grid with its own cell lists (2nx*2ny*2nz). */
int orc_mc_hash_sort(orc_state* s);
/* Node field of the last particle-source mesh ((2nx+1)(2ny+1)(2nz+1) doubles)
   and the 2N cell lists (either pointer may be NULL). */
int orc_get_mc_field(orc_state* s, double* out);
int orc_get_mc_cells(orc_state* s, uint32_t* start, uint32_t* end);
int orc_apply_forces(orc_state* s);
int orc_levelset_update(orc_state* s);
int orc_p2g(orc_state* s);
int orc_extrapolate_bc(orc_state* s, int recompute_phi);
int orc_project(orc_state* s, mpf_solve_info* info);
int orc_g2p(orc_state* s);
int orc_advect(orc_state* s);
int orc_heat(orc_state* s, mpf_solve_info* info);
/* The two halves of orc_heat after the particle-to-grid map: the grid step on
   the state's TEMPERATURE (heater, implicit diffusion, boundary copy,
   TEMPERATURE_CHANGE), and the grid-to-particle map with latent heat. */
int orc_heat_grid(orc_state* s, mpf_solve_info* info);
int orc_heat_g2p(orc_state* s);
/* Single pieces of orc_levelset_update on the state's fields: the level-set
   gradient of LSMERGE, and the count-based and solid cell flags (MARK) from
   the current cell lists. */
/* Pressure solve with FP32 vectors (the reference's arithmetic; dot
   products still summed in double): the CPU baseline's form. */
int orc_set_solve_float(orc_state* s, int on);
int orc_phigrad(orc_state* s);
int orc_mark_from_counts(orc_state* s);
int orc_mark_solid_cells(orc_state* s);
int orc_phase_change(orc_state* s, int frame, mpf_phase_info* info);
int orc_marching_cubes(orc_state* s, mpf_mc_info* info);
int orc_get_mesh(orc_state* s, double* verts, uint32_t* tris);
int orc_get_mc_cases(orc_state* s, uint8_t* cases);
int orc_step(orc_state* s, int frame, mpf_step_info* info);

/* Pieces exposed for closed-form tests. */
uint32_t orc_hash_key(const mpf_config* cfg, float x, float y, float z);
void orc_stable_sort_pairs(int64_t n, uint32_t* keys, uint32_t* vals);
/* Apply the pressure operator to x (double in/out) using the state's flags/ls. */
int orc_apply_operator(orc_state* s, const double* x, double* y);
int orc_apply_heat_operator(orc_state* s, const double* x, double* y);
double orc_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);
int orc_mc_tables(int32_t* edge_table, int32_t* tri_table, int32_t* num_verts);

#ifdef __cplusplus
}
#endif
#endif
