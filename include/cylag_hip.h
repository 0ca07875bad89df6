/*
 * cylag_hip.h -- C ABI of the MI355X particle time step (libcylag_hip.so).
 *
 * The reference has no native boundary: its per-particle kernels are Cython
 * cdef functions reachable only from Solver.forward (cylag/core/solver.pyx:413,
 * a cpdef).  The replaceable unit is therefore the whole step.  These entry
 * points are what the Python layer (cylag_amd, via ctypes) binds in place of
 * the Cython modules; each cites the reference code it replaces.
 *
 * Conventions: every call returns 0 on success or a negative CYL_E* code.
 * Caller-owned host buffers are copied; no pointer is kept after return.  The
 * context owns all device memory.  A context is used by one host thread.
 * Arrays are C-contiguous in the reference's shapes (AoS (npart, dim) for
 * particle rows, (ntimes, npoints*nlayers) layer-major for fields).
 */
#ifndef CYLAG_HIP_H
#define CYLAG_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CYL_OK 0
#define CYL_E_ARG -1
#define CYL_E_HIP -2
#define CYL_E_STATE -3
#define CYL_E_STEP -4 /* device-detected condition: see cyl_last_error */

typedef struct cyl_ctx cyl_ctx;

/* Solver parameters (cylag/cylag_dico.py:9-51, resolved by parameters.pyx:116-195). */
typedef struct {
  int model, time_scheme, particle_velocity_init, rng_method;
  int diffusion_model, diffusion_lsm2_option;
  double horizontal_diffusivity, vertical_diffusivity, schmidt_number;
  int boundary_conditions, boundary_conditions_type, max_wall_rebounds, bnd_statistics;
  double rebound_damping_coef, wall_roughness_angle, stranding_probability;
  int compute_depth;
  int frozen;
  int lsm3_edge_opt;                                                  /* 1 series, 2 clipped closed form */
  double diffusion_lsm3_tl_horizontal, diffusion_lsm3_tl_vertical;  /* TL of the constant-K models */
  int stranding_output;                                               /* record stranded positions */
  int bottom_deposition; /* 1: a particle below the bed is deposited (recorded, then deleted) instead of
                            reflected; an extension with no counterpart in the reference (0 = its
                            damped reflection, compute_bc.pyx:294-332) */
} cyl_params;

/* Particle-set constants (lagrangian_particle_set.pyx:158-261,349-375). */
typedef struct {
  double rho0, density, drag_coef, added_mass_coef, buoyancy_velocity_value;
  int drag_coef_model, added_mass_force, buoyancy_velocity_model;
  double a0c, a1c, a2c, a3c;
} cyl_pconst;

/* ---- context ---------------------------------------------------------- */
int cyl_ctx_create(int device, int flags, cyl_ctx **out);
int cyl_ctx_destroy(cyl_ctx *ctx);
/* Code, first failing pool index and message of the last failed call
 * (replaces the ValueErrors raised inside the step, e.g.
 * eulerian_field_set.pyx:1092,1272-1276). */
int cyl_last_error(cyl_ctx *ctx, int *code, int *particle, char *msg, size_t n);
int cyl_device_info(cyl_ctx *ctx, char *name, size_t n, size_t *free_bytes, size_t *total_bytes);

/* ---- mesh: TriangularMesh tables (triangular_mesh.pxd:3-45) ------------- */
int cyl_mesh_upload(cyl_ctx *ctx, int nv, const double *x, const double *y, int nt, const int *tri,
                    int ne, const int *edges, const int *edgetri, const int *triedge,
                    const int *edge_labels, const double *det, const double *gradx,
                    const double *grady, int nopen);

/* ---- fields: EulerianFieldSet (eulerian_field_set.pyx:125-256) ---------- *
 * fids = {fid_zs, fid_ux, fid_uy, fid_uz, fid_turbeng, fid_dissip, fid_zb}.
 * velocity_z may be NULL in 2D, opt_fields NULL when nopt == 0. */
int cyl_fields_upload(cyl_ctx *ctx, int dim, int nlayers, int npoints, int ntimes,
                      const double *times, const double *elevation, const double *velocity_x,
                      const double *velocity_y, const double *velocity_z, int nopt,
                      const double *opt_fields, const int *fids, int friction_model,
                      double friction_coefficient);
/* Set the frozen flag and reset time interpolation state
 * (solver.pyx:140-155: record from initial time, _interp_done = _prev_valid = False). */
int cyl_fields_reset(cyl_ctx *ctx, int frozen, double initial_time);
/* _time_interpolation(time) (eulerian_field_set.pyx:866-938) on the device. */
int cyl_fields_interpolate(cyl_ctx *ctx, double time);
/* Download a time-interpolated node array (field id, all layers). */
int cyl_fields_download_tmp(cyl_ctx *ctx, int fid, double *out);

/* ---- particles: LagrangianParticleSet arrays in pool order ------------- *
 * position/velocity/fluid_velocity_seen are (npart, dim); init_tri holds for
 * tri == -2 rows the element found on the host by a trifinder (-3: search on
 * the device, the reference's KD-tree path). */
int cyl_particles_upload(cyl_ctx *ctx, int npart, int dim, const double *position,
                         const double *velocity, const double *fluid_velocity_seen,
                         const int *inactive, const int *tri, const int *lowerlayer,
                         const double *diameter, const double *volume, const int *init_tri,
                         const cyl_pconst *pconst);
/* Any of the outputs may be NULL.  last_* are the rows at the start of the
 * last step (lagrangian_particle_set.pyx:237-240 semantics). */
int cyl_particles_download(cyl_ctx *ctx, int npart, double *position, double *velocity,
                           double *fluid_velocity_seen, int *inactive, int *tri, int *lowerlayer,
                           double *last_position, double *last_velocity, int *last_tri,
                           double *depth);
/* Rows deleted by the last step, in pool order (for tag renumbering,
 * lagrangian_particle_set.pyx:786-788); returns their count in *n. */
int cyl_particles_deleted(cyl_ctx *ctx, int *pool_index, int cap, int *n);
/* Rows deleted since the particles were uploaded (device counter). */
int cyl_deleted_count(cyl_ctx *ctx, long long *n);
/* Write new rows into pool slots (LagrangianParticleSet.add first fit,
 * lagrangian_particle_set.pyx:680-710); rows get tri = -2. */
int cyl_particles_add(cyl_ctx *ctx, int n, const int *slots, const double *position,
                      const double *diameter, const double *volume, const int *init_tri);

/* ---- random stream: random_utils.pyx:17-53 ----------------------------- */
int cyl_rng_set(cyl_ctx *ctx, uint64_t state, uint64_t inc);
int cyl_rng_get(cyl_ctx *ctx, uint64_t *state, uint64_t *inc);
/* Host-side stream helpers used by the module-level API (random_utils.pxd:1-12). */
uint32_t cyl_pcg32_next(uint64_t *state, uint64_t inc);
void cyl_pcg32_seed(uint64_t *state, uint64_t *inc, uint64_t seed, uint64_t stream);
void cyl_pcg32_advance(uint64_t *state, uint64_t inc, uint64_t delta);
double cyl_random_normal(uint64_t *state, uint64_t inc, int method);
void cyl_random_pair(uint64_t *state, uint64_t inc, int method, double *a, double *b);

/* ---- the step ----------------------------------------------------------- *
 * initialize_particle_state (solver.pyx:238-380) for every tri == -2 row. */
int cyl_init_particles(cyl_ctx *ctx, const cyl_params *par, double time, double dt);
/* nsteps x Solver.forward(dt) (solver.pyx:413-585: time interpolation, pool
 * loop with LSM update and boundary conditions, time increment).  bnd_stats,
 * if not NULL, receives nsteps*nopen open-boundary exit counts.  *time_inout
 * is advanced exactly like SimuTime.increment (simutime.pyx:77-85). */
int cyl_step(cyl_ctx *ctx, const cyl_params *par, double *time_inout, double dt, int nsteps,
             int *bnd_stats);
/* ---- batched scenarios ------------------------------------------------------ *
 * No counterpart in the reference (the nearest is SolverParallel's split of one
 * release into pool-index ranges with per-rank streams, solver_parallel.pyx:99-130).
 * The uploaded pool is split into nscen contiguous pool ranges [base[s], base[s+1]);
 * scenario s has its own parameters, random stream (state, inc) and number of steps;
 * all share the mesh, the fields, the time step, the particle constants and one
 * rng_method.  Each scenario's result is the one a plain run of it alone gives. */
int cyl_scenarios_set(cyl_ctx *ctx, int nscen, const int *base, const cyl_params *par,
                      const uint64_t *state, const uint64_t *inc, const int *nsteps);
/* Random stream of scenario s after the steps run so far. */
int cyl_scenario_rng_get(cyl_ctx *ctx, int s, uint64_t *state, uint64_t *inc);
/* nsteps steps of every scenario that has steps left (finished scenarios keep
 * their rows).  bnd_stats, if not NULL, receives nsteps*nscen*nopen counts. */
int cyl_batch_step(cyl_ctx *ctx, double *time_inout, double dt, int nsteps, int *bnd_stats);
/* Continuous batching (no counterpart in the reference; it serves callers that
 * keep many releases resident on one GPU and let them join and leave).
 * cyl_batch_collect: the rows of scenario s in the order of its own pool, as
 * cyl_particles_download gives them, and its deleted rows in deletion order
 * (lagrangian_particle_set.pyx:786-788; reported rows become plain inactive
 * rows).  Arrays may be NULL. */
int cyl_batch_collect(cyl_ctx *ctx, int s, int npart, double *pos, double *vel, double *us, int *inactive,
                      int *tri, int *lowerlayer, int *deleted, int cap, int *ndeleted);
/* Allocation-free variant of cyl_batch_collect, same arguments and results.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree: the rows are staged in a
 * static device array that each GPU the library runs on holds until the
 * process ends (CYL_NOALLOC_MAX_ROWS * 92 bytes = 46 MiB); scenarios with more
 * rows are rejected with CYL_E_ARG.  Calls share that array and a host buffer
 * of deletion keys, so the library serialises them with a process-wide lock. */
#define CYL_NOALLOC_MAX_ROWS 524288
int cyl_batch_collect_noalloc(cyl_ctx *ctx, int s, int npart, double *pos, double *vel, double *us,
                              int *inactive, int *tri, int *lowerlayer, int *deleted, int cap, int *ndeleted);
/* A new batch on the same device state: the nkeep scenarios keep[] (increasing
 * ids) keep their rows, random streams and remaining steps and become
 * scenarios 0..nkeep-1; nnew scenarios follow, with their rows given as in
 * cyl_particles_upload (concatenated; rows with tri = -2 are initialized at
 * `time`, as Solver's constructor does, solver.pyx:193-197), parameters,
 * streams and step counts.  The other scenarios leave the device.  The new
 * scenarios' particle constants are the context's. */
int cyl_batch_reshape(cyl_ctx *ctx, int nkeep, const int *keep, int nnew, const int *new_npart,
                      const cyl_params *new_par, const uint64_t *new_state, const uint64_t *new_inc,
                      const int *new_nsteps, const double *pos, const double *vel, const double *us,
                      const int *inactive, const int *tri, const int *lowerlayer, const double *diameter,
                      const double *volume, const int *init_tri, double time, double dt);

/* Tuning (no counterpart in the reference): element re-sort interval in steps
 * (0 = never; default 8, and the first step after an upload re-sorts), the
 * resident 256-thread blocks per CU of the persistent grids (0 = as many as
 * fit), where the 2D fused kernels take their normals from (0: a generator
 * that draws runs of ranks in pool order, the default; 1: sampled inline in
 * the kernel; 2: a generator that draws each slot's pair at its rank's stream
 * position and stores it in slot order), and whether multi-step calls replay
 * a captured HIP graph of 2 * sort_interval steps (uniform draw assignment,
 * no boundary statistics; otherwise the steps are enqueued one by one).
 * Results do not depend on any of them. */
int cyl_set_tuning(cyl_ctx *ctx, int sort_interval, int blocks_per_cu, int inline_normals, int use_graph);
/* Timing of the last cyl_step: summed duration of the fused step-kernel
 * launches (HIP events on the context's stream), wall time of the call, and
 * the number of timed launches. */
int cyl_last_timing(cyl_ctx *ctx, double *step_kernel_ms, double *total_ms, int *kernel_launches);
/* Control sections (solver.pyx:540-549, control_shapes.pyx:56-85): section s
 * owns nseg_of[s] consecutive segments (x0, y0, x1, y1) of seg[].  Each step
 * counts, per section, the rows active at its start whose path [last
 * position, position] meets one of the section's segments (nsec = 0 clears). */
int cyl_control_sections_set(cyl_ctx *ctx, int nsec, const int *nseg_of, const double *seg);
/* Counts of the last cyl_step call, [nsteps][nsec] (nout = nsteps * nsec). */
int cyl_control_counts(cyl_ctx *ctx, int *out, int cap, int *nout);
/* Stranded positions of the last cyl_step call (stranding_output; compute_bc.pyx:107-122):
 * n records of (x, y, z) in the reference's order (steps in turn, rows in pool order). */
int cyl_stranded_get(cyl_ctx *ctx, double *xyz, int cap, int *n);
/* Deposits of the last cyl_step call (bottom_deposition; no counterpart in the reference): for each
 * particle that reached the bed, its position on the bed (x, y, bed elevation) and the step of the
 * call (0-based) in which it deposited, in the serial order of the reference's loop (steps in
 * turn, rows in pool order).  n receives the count; at most cap records are copied.  Plain
 * (single-scenario) contexts only. */
int cyl_deposited_get(cyl_ctx *ctx, double *xyz, int *step, int cap, int *n);
/* Which kernels advanced the particles in the last cyl_step / cyl_batch_step
 * call (no counterpart in the reference): launches of the fused kernels and of
 * the generic kernel that did the step (guarded reruns that found nothing to
 * do are not counted; steps inside graph replays are not counted), and the
 * number of steps run by graph replays. */
int cyl_last_path(cyl_ctx *ctx, int *fused_launches, int *generic_launches, int *graph_steps);
/* Summed duration of what prepares the fused kernels' normals in the timed
 * steps of the last cyl_step: the normal generator, or the draw-start pass of
 * the inline samplers (always in 3D); 0 when the generic kernel ran. */
int cyl_last_prep_timing(cyl_ctx *ctx, double *prep_ms);

#ifdef __cplusplus
}
#endif
#endif
