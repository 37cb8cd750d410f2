/*
 * exocubed.h -- C ABI of the MI355X (gfx950) cubed-sphere hydro library.
 *
 * One handle holds one model instance (six panels, one MeshBlock per panel)
 * resident in HBM and binds one HIP stream.  No C++ exception crosses this
 * boundary: every call returns EXO_OK or a negative EXO_ERR_* code, and
 * exo_last_error() returns the message of the most recent failure on a handle.
 * Calls on different handles may run concurrently; calls on one handle must
 * not overlap.
 *
 * Host arrays use the reference's MeshBlock layout: for each panel p (block
 * ID p+1 of the 2 x 3 panel grid, src/exo3/cs_find_block_id.cpp:11-75), an
 * Athena array (NVAR, nc3, nc2, nc1) with x1 (radial) fastest, NGHOST = 3,
 * nc1 = nx1 + 6 and nc2 = nc3 = nx + 6.  Panels follow each other: element
 * (p, n, k, j, i) sits at (((p*NVAR + n)*nc3 + k)*nc2 + j)*nc1 + i.
 * NVAR is 4 for the shallow-water EOS (IDN = g h, IVX = 0, IVY, IVZ) and 5 for
 * the ideal-gas EOS (IDN, IM1, IM2, IM3, IEN).
 */
#ifndef EXOCUBED_H_
#define EXOCUBED_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the library is built with hidden visibility; only these entry points export */
#ifndef EXO_API
#define EXO_API __attribute__((visibility("default")))
#endif

#define EXO_OK 0
#define EXO_ERR_ARG (-1)
#define EXO_ERR_HIP (-2)
#define EXO_ERR_STATE (-3)
#define EXO_ERR_NOMEM (-4)
#define EXO_ERR_UNSUPPORTED (-5)
#define EXO_ERR_IO (-6)

/* equation of state: cmake/exo2.cmake:14 (shallow_yz), parameters.cmake:18 */
enum { EXO_EOS_SHALLOW_YZ = 0, EXO_EOS_IDEAL = 1 };
/* Riemann solver: src/snap/riemann/{roe_shallow_yz,lmars_shallow_yz,
 * hllc_transform,lmars}.cpp */
enum { EXO_RS_ROE_SW = 0, EXO_RS_LMARS_SW = 1, EXO_RS_HLLC = 2, EXO_RS_LMARS = 3 };
/* Coriolis source: none, W92.cpp:48-70 Forcing2, hs94.cpp rotation term */
enum { EXO_COR_NONE = 0, EXO_COR_FORCING2 = 1, EXO_COR_3D = 2 };
/* arithmetic of the explicit flux path */
enum { EXO_PREC_FP64 = 0, EXO_PREC_MIXED = 1 };

typedef struct exo_config {
  int nx;            /* cells per panel edge (even) */
  int nx1;           /* radial cells */
  int eos;           /* EXO_EOS_* */
  int rsolver;       /* EXO_RS_* */
  int implicit_flag; /* <hydro> implicit_flag (solve_implicit_3d.cpp:102-135): bit 1 solves
                        the x1 columns, with the full 5x5 system when bit 8 is set
                        (full_correction.cpp), else the partial 3x3 one
                        (partial_correction.cpp); 0 and 8: u += du.  The x2/x3 bits
                        (2, 4) are rejected. */
  int ntracer;       /* passive scalars (upwinded by the mass flux) */
  int coriolis;      /* EXO_COR_* */
  int cfl_x1;        /* 1: radial widths enter the CFL minimum */
  double x1min, x1max;
  double gamma;
  double grav;       /* <hydro> grav_acc1: negative = downward; 0 for shallow water */
  double omega;      /* planetary rotation rate for the Coriolis source */
  double cfl;        /* <time> cfl_number */
  int precision;     /* EXO_PREC_* */
} exo_config;

typedef struct exo_handle exo_handle;

/* --- lifetime ---------------------------------------------------------- */
/* Replaces the per-MeshBlock construction of GnomonicEquiangle
 * (src/exo3/gnomonic_equiangle.cpp:24-221), CubedSphere (cubed_sphere.cpp:25-35),
 * Decomposition (decomposition.cpp:25-48) and ImplicitSolver
 * (implicit_solver.cpp:34-90) for all six panels at once. */
EXO_API int exo_create(const exo_config *cfg, exo_handle **out);
EXO_API int exo_destroy(exo_handle *h);
EXO_API const char *exo_last_error(const exo_handle *h);
/* number of conserved variables (NHYDRO) and padded cells per panel */
EXO_API int exo_nvar(const exo_handle *h);
EXO_API size_t exo_state_doubles(const exo_handle *h); /* 6*NVAR*nc3*nc2*nc1 */

/* --- state transfer (host layout above) -------------------------------- */
/* Upload conserved variables u and (if ntracer > 0) conserved scalars s = rho r
 * (NTRACER, nc3, nc2, nc1 per panel; may be NULL).  Interior cells are read;
 * ghosts are then filled as Mesh::Initialize does on the cube (ghost exchange,
 * patches/29.mesh.patch:19-86, then cons->prim and the reflecting x1 BC). */
EXO_API int exo_upload_state(exo_handle *h, const double *u, const double *s);
EXO_API int exo_download_state(exo_handle *h, double *u, double *s);
EXO_API int exo_download_prim(exo_handle *h, double *w);

/* --- the hot path, one call per reference entry point ------------------- */
/* Hydro::CalculateFluxes (patches/19.calculate_fluxes.patch:17-163): WENO5
 * reconstruction (Reconstruction::Weno5X1/X2/X3), hydrostatic decomposition
 * around the x1 sweep (change_to_perturbation.cpp:39-190), panel-edge L/R
 * exchange (cubed_sphere.cpp:267-687) and Hydro::RiemannSolver on every face. */
EXO_API int exo_calc_fluxes(exo_handle *h);
/* Copy the face fluxes of direction dir (0 = x1, 1 = x2, 2 = x3) to the host in
 * Athena's flux layout: x1 (NVAR, nc3, nc2, nc1+1), x2 (NVAR, nc3, nc2+1, nc1),
 * x3 (NVAR, nc3+1, nc2, nc1) per panel; only interior faces are defined. */
EXO_API int exo_download_flux(exo_handle *h, int dir, double *out);
/* Hydro::AddFluxDivergence(wght, u) followed by
 * GnomonicEquiangle::AddCoordTermsDivergence (gnomonic_equiangle.cpp:834-879)
 * into u (shallow water) or into the tendency du (ideal gas, zeroed first, as
 * ImplicitHydroTasks::IntegrateHydro, implicit_hydro_tasks.cpp:157-162).
 * It applies the first SSPRK3 stage's register weights (u1 <- u, u kept): at
 * stages 2 and 3 use exo_integrate_hydro, which forms that stage's average. */
EXO_API int exo_add_flux_divergence(exo_handle *h, double wght);
/* ImplicitHydroTasks::IntegrateHydro (implicit_hydro_tasks.cpp:128-186) with
 * AddSourceTerms (:188-214) at SSPRK3 stage 1..3: the stage's weighted average
 * of u and the register u1 (SSPRK3 Shu-Osher weights), then
 * AddFluxDivergence + AddCoordTermsDivergence + sources with
 * wght = beta(stage) dt into du (ideal gas) or u (shallow water).  Followed by
 * exo_implicit_correct(h, beta(stage) dt) and exo_fill_ghosts it gives the
 * same bits as exo_stage(h, stage, dt). */
EXO_API int exo_integrate_hydro(exo_handle *h, int stage, double dt);
/* CubedSphereUtility::PackData panel-edge remap (cubed_sphere_utility.cpp:
 * 371-617) of the conserved state into the face ghost zones, then
 * cons->prim (EquationOfState::ConservedToPrimitive) and the reflecting x1 BC. */
EXO_API int exo_fill_ghosts(exo_handle *h);
/* ImplicitSolver::SolveImplicit3D (solve_implicit_3d.cpp:27-180) on the x1
 * columns, FullCorrection (implicit_flag 9, full_correction.cpp) or
 * PartialCorrection (implicit_flag 1, partial_correction.cpp:26-177), applied
 * to du, then u += du (ImplicitHydroTasks::ImplicitCorrection,
 * implicit_hydro_tasks.cpp:245-266). */
EXO_API int exo_implicit_correct(exo_handle *h, double dt_stage);
/* Hydro::NewBlockTimeStep (patches/21.new_blockdt.patch:16-37) and
 * Mesh::NewTimeStep's factor-2 growth limit; returns the next dt. */
EXO_API int exo_new_dt(exo_handle *h, double *dt);
/* One SSPRK3 stage (1..3) with time step dt, as the task list runs it. */
EXO_API int exo_stage(exo_handle *h, int stage, double dt);
/* nsteps full SSPRK3 steps, dt from the CFL condition on the device when
 * dt_fixed <= 0; *dt_last (may be NULL) receives the last step's dt.
 * Replaces the main loop body (main.cpp.in:130-186). */
EXO_API int exo_step(exo_handle *h, int nsteps, double dt_fixed, double *dt_last);
/* Block until all work queued on the handle's stream has finished. */
EXO_API int exo_sync(exo_handle *h);
/* Live kernel timing: between begin and end every launch on the handle's stream
 * is bracketed by a HIP event pair; end returns the summed milliseconds and
 * launch counts per kernel id (0 prim, 1 ghost, 2 flux_x1, 3 edge_states,
 * 4 flux_x2, 5 flux_x3, 6 update, 7 implicit, 8 dt, 9 output). */
EXO_API int exo_profile_begin(exo_handle *h);
EXO_API int exo_profile_end(exo_handle *h, double *ms, int *counts, int nkinds);

/* --- conserved totals and launch tuning ---------------------------------- */
/* Sum over the interior cells of u_n * CellVolume for every conserved
 * variable (out5[n], n < NVAR; the rest 0) in FP64 with compensated
 * (TwoSum) addition in a fixed order: deterministic, accurate to about one
 * rounding of each total.  The reference has no such diagnostic; it checks the
 * conservation of the finite-volume update (mass, and energy without sources). */
EXO_API int exo_conserved_sums(exo_handle *h, double *out5);
/* Autotune the launches of the handle's kernels that only read the conserved
 * state or only write the tendency (repeated launches change nothing): the
 * persistent launch width (the face kernels, the fused x2-face stage update
 * of instances without passive scalars, the primitive conversion, the ghost
 * remap, the CFL minimum), and the x1 sweep's block shape (waves per
 * 64-column block: 2, 4 or 8).
 * Each candidate is timed `reps` times and the fastest kept; a configuration
 * tuned before in this process reuses its result without timing.  caps (may
 * be NULL) receives up to ncaps values: the width of each launch kind (0 = one
 * workgroup per block), then, from index EXO_NKINDS on, its shape (0 = the
 * default). */
#define EXO_NKINDS 30
EXO_API int exo_tune(exo_handle *h, int reps, int *caps, int ncaps);

/* Device bytes exo_create allocates for this configuration (host only; 0 for
 * an invalid configuration).  Used to admit work by memory footprint.  Not
 * included: what a handle allocates on first use of exo_output_fields /
 * exo_write_netcdf (the float fields of one output, plus the "uov" factors) and
 * of exo_save_states / exo_step_redo (one copy of u, w and the scalars). */
EXO_API size_t exo_footprint(const exo_config *cfg);

/* --- batches of independent instances ----------------------------------- */
/* A batch advances several handles (uneven sizes, any configuration) with one
 * launch per phase over all of them: each kernel's blocks are split over the
 * members by a prefix table and a launch may be persistent (fewer workgroups
 * than blocks, looping).  With EXO_BATCH_GRAPH each SSPRK3 step is captured
 * once as a HIP graph and replayed.  Every member keeps its own state, dt and
 * CFL; results are identical to stepping each member alone.  Members must be
 * idle (no call in flight) while a batch call runs, and must outlive the
 * batch.  No reference counterpart: the reference runs one instance per
 * process (main.cpp.in:130-186). */
typedef struct exo_batch exo_batch;
#define EXO_BATCH_GRAPH 1
EXO_API int exo_batch_create(exo_handle *const *members, int n, int flags, exo_batch **out);
EXO_API int exo_batch_destroy(exo_batch *b);
EXO_API const char *exo_batch_last_error(const exo_batch *b);
/* nsteps SSPRK3 steps of every member (dt from each member's CFL when
 * dt_fixed <= 0); returns when the work has finished */
EXO_API int exo_batch_step(exo_batch *b, int nsteps, double dt_fixed);
/* exo_conserved_sums of every member: out[member][5] */
EXO_API int exo_batch_sums(exo_batch *b, double *out);
/* exo_tune over the batch's launches (reused for a batch of the same member
 * configurations); exo_batch_set_caps sets widths, and from index EXO_NKINDS
 * on shapes, directly */
EXO_API int exo_batch_tune(exo_batch *b, int reps, int *caps, int ncaps);
EXO_API int exo_batch_set_caps(exo_batch *b, const int *caps, int ncaps);
/* live kernel timing of the batch's launches (kernel ids as exo_profile_end;
 * graphs are not used while profiling) */
EXO_API int exo_batch_profile_begin(exo_batch *b);
EXO_API int exo_batch_profile_end(exo_batch *b, double *ms, int *counts, int nkinds);
EXO_API int exo_batch_sync(exo_batch *b);

/* --- serving: continuous batching of forward-run requests ---------------- */
/* A request loop in the library: prep threads build initial states, one
 * scheduler thread admits prepared requests in arrival order while their
 * device footprint (exo_footprint) fits mem_budget bytes (0: 80 % of the free
 * device memory) and fewer than max_active run, keeps their state resident in
 * HBM and advances all active requests together through one exo_batch in
 * chunks of at most chunk_steps steps (flags: EXO_BATCH_GRAPH).  Requests that
 * arrive while others run join at the next chunk boundary; finished requests
 * free their memory at once.  A request's result does not depend on the mix.
 * No reference counterpart: the reference runs one instance per process. */
typedef struct exo_server exo_server;
EXO_API int exo_server_create(size_t mem_budget, int max_active, int chunk_steps, int flags,
                      int prep_threads, exo_server **out);
/* Queue a forward run of nsteps SSPRK3 steps (CFL dt).  The initial state is
 * copied from u0/s0 (host layout above), or, with u0 == NULL, built by
 * exo_init_state_cfg(cfg, kind, seed, params).  keep_state = 0 drops the final
 * state (only the info of exo_server_wait is kept).  Fails with EXO_ERR_NOMEM
 * if the request can never fit the budget. */
EXO_API int exo_server_submit(exo_server *s, const exo_config *cfg, const double *u0, const double *s0,
                      int kind, uint64_t seed, const double *params, int nparams, int nsteps,
                      int keep_state, uint64_t *id_out);
/* Wait (timeout_s <= 0: forever) for a request, copy its final state to
 * u_out/s_out (may be NULL) and release it.  info[10] (may be NULL): latency,
 * queue wait, run time (s), last dt, cells, steps, mass before/after, energy
 * (shallow water: x2 momentum) total before/after. */
EXO_API int exo_server_wait(exo_server *s, uint64_t id, double timeout_s, double *u_out, double *s_out,
                    double *info);
/* stats[9]: completed, cell-updates, seconds first submit to last completion,
 * seconds in batch steps, batches formed, chunks, peak active, peak bytes,
 * budget bytes */
EXO_API int exo_server_stats(exo_server *s, double *stats, int n);
EXO_API const char *exo_server_last_error(const exo_server *s);
/* Finishes the requests already on the device, drops queued ones. */
EXO_API int exo_server_destroy(exo_server *s);

/* --- problem set-up (synthetic initial states of the benchmark configs) - */
/* Fill host conserved arrays (interior cells) for a seeded initial state.
 * kind 0: shallow-water random-mode height field in geostrophic balance
 *         (params: mean depth, mode amplitude, g);
 * kind 1: isothermal hydrostatic atmosphere with Gaussian temperature noise
 *         and random-phase zonal waves (params: T0, ps, Rd, noise fraction,
 *         wave amplitude);
 * kind 2: lapse-rate atmosphere to an isothermal stratosphere (params: Ts,
 *         lapse K/m, Tstrat, ps, Rd, noise K). */
EXO_API int exo_init_state(exo_handle *h, int kind, uint64_t seed, const double *params,
                   int nparams, double *u_out, double *s_out);
/* The same without a handle (host only, no GPU needed). */
EXO_API int exo_init_state_cfg(const exo_config *cfg, int kind, uint64_t seed, const double *params,
                       int nparams, double *u_out, double *s_out);

/* --- single-face and single-column shims for per-function parity ------- */
/* Riemann flux of one face with the given L/R primitive states, computed by
 * the device kernel (dir 0/1/2 = x1/x2/x3; the metric of face (p,k,j) of this
 * handle's grid is used). */
EXO_API int exo_riemann_face(exo_handle *h, int dir, int p, int k, int j, const double *wl,
                     const double *wr, double *flx);
/* ImplicitSolver::FullCorrection on one x1 column (full_correction.cpp:32-165
 * with forward_backward.hpp:15-126), computed by the device's column solve:
 * w_col (5, nx1 + 6) primitives of the column with its x1 ghosts, velocities
 * already projected as SolveImplicit3D projects them (solve_implicit_3d.cpp:
 * 117-123); du_col (5, nx1 + 6) the tendency, its interior levels replaced by
 * the correction; area1 the nx1 + 1 interior x1 face areas, vol the nx1
 * volumes; gamma and grav of this handle's configuration (ideal gas). */
EXO_API int exo_implicit_column(exo_handle *h, int nx1, const double *w_col, double *du_col,
                                const double *area1, const double *vol, double dt);
/* No counterpart in the reference: exo_implicit_column for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  ncol columns at once, all arrays in
 * device memory of the handle's GPU: column c's w and du at c*5*(nx1 + 6),
 * its areas at c*(nx1 + 1), its volumes at c*nx1; du's interior levels are
 * replaced by the correction.  The coefficient scratch is a fixed device
 * array of EXO_NOALLOC_SCRATCH_DOUBLES doubles per GPU, 64-column blocks of
 * nx1*1920 doubles each: larger requests are rejected (EXO_ERR_ARG).  Nothing
 * is allocated and nothing synchronises: the work is queued on the handle's
 * stream (exo_sync waits for it).  Calls that overlap on one GPU share the
 * scratch and race. */
#define EXO_NOALLOC_SCRATCH_DOUBLES (1 << 19)
EXO_API int exo_implicit_columns_noalloc(exo_handle *h, int ncol, int nx1, const double *w_cols,
                                         double *du_cols, const double *area1, const double *vol,
                                         double dt);
/* WENO5 value at the face between b and c seen from cell c (device). */
EXO_API int exo_weno5_batch(exo_handle *h, const double *in5, double *out, int n);
/* deterministic device exp/log/pow on n inputs (op 0 exp, 1 log, 2 pow y),
 * (op 3) the shared-reciprocal division x[2t] / x[2t+1] the implicit solve
 * uses, or (op 4) the device's unscaled IEEE division sequence of x[2t] / x[2t+1]
 * that the FP64 kernels use for every quotient */
EXO_API int exo_math_batch(exo_handle *h, int op, const double *x, double y, double *out, int n);

/* Host-only introspection of the set-up tables (no GPU needed): which 0 ghost
 * remap descriptors, 1 edge-exchange descriptors, 2 metric tables
 * [20][6][nc3][nc2], 3 x1f|x1v|dx1f, 4 x2f and 5 x3f of the six panels,
 * 6 per-row tangent tables [8][6][nc2 + 1] (x = tan x2v, tan x2f, C =
 * sqrt(1 + x^2) at centres and faces, then the same for y along x3).
 * With buf == NULL only *count is returned. */
EXO_API int exo_describe(const exo_config *cfg, int which, void *buf, size_t cap_bytes,
                 size_t *count);

/* --- output: the reference's per-MeshBlock NetCDF files ------------------ */
/* Replaces NetcdfOutput::WriteOutputFile (src/outputs/netcdf.cpp:34-469) for
 * the output variables this path has:
 *   "prim": rho, press (ideal gas), vel1, vel2, vel3, r0..r{NTRACER-1}
 *   "cons": dens, Etot (ideal gas), mom1, mom2, mom3, s0..s{NTRACER-1}
 *   "uov" : lat, lon, U, V, sqrtg, the user outputs of
 *           examples/2023-Chen-exo3/W92.cpp:115-151 (MeshBlock::
 *           UserWorkBeforeOutput; degrees, lat-lon wind, delta / (C D))
 * Field names in file order, comma separated (host only, no GPU needed). */
EXO_API int exo_output_names(const exo_config *cfg, const char *variable, char *names,
                             size_t cap, int *nvar);
/* The fields of one output variable of the current state as float, laid out
 * as the files hold them: [6 panels][nvar][n1][n2][n3], x3 fastest, n the
 * interior cells or (ghost_zones != 0) the cells with three ghosts on each
 * side.  Computed on the device from the FP64 state of the last full stage. */
EXO_API int exo_output_fields(exo_handle *h, const char *variable, int ghost_zones, float *out,
                              size_t cap_floats);
/* Writes <basename>.block<gid>.<file_id>.<file_number %05d>.nc for every
 * MeshBlock: each panel split into blocks_per_side^2 blocks (a power of two
 * dividing N), gid in Athena's MeshBlockTree order (panel, then Z-order of
 * the block within it).  file_id NULL = the variable name (Athena's default
 * "id").  NetCDF-3 64-bit-offset files (the reference asks libnetcdf for
 * NetCDF-4); x3f holds x3 face coordinates (the reference copies x1f there,
 * netcdf.cpp:306-310). */
EXO_API int exo_write_netcdf(exo_handle *h, const char *basename, const char *variable,
                             const char *file_id, int file_number, double time,
                             int blocks_per_side, int ghost_zones);
/* The same files from host fields in the exo_output_fields layout (host only,
 * no GPU needed); EXO_ERR_IO when a file cannot be written. */
EXO_API int exo_write_netcdf_fields(const exo_config *cfg, const char *basename,
                                    const char *variable, const char *file_id, int file_number,
                                    double time, int blocks_per_side, int ghost_zones,
                                    const float *fields);

/* --- failure recovery: redo a step with half the time step ---------------- */
/* The reference's main loop (main.cpp.in:126-160) keeps a copy of the state
 * before each step when max_redo > 0 (Mesh::SaveAllStates,
 * src/mesh_redo.cpp:12-17), and after the step checks that every block's
 * primitive conversion found no negative or NaN density or pressure
 * (ideal_moist_hydro.cpp:83,108,150; Mesh::CheckAllValid, mesh_redo.cpp:31-45).
 * An invalid step is undone (LoadAllStates) and repeated with half the time
 * step (DecreaseTimeStep); more than max_redo + 1 redone steps over the run
 * fail with "Too many attempts to redo the step".
 *   exo_state_valid: 1 if the last stage's primitive conversion (or the
 *     upload's) found every density and pressure non-negative, else 0; the
 *     shallow-water EOS never marks a state invalid.
 *   exo_save_states / exo_load_states: snapshot of the conserved, primitive
 *     and scalar state and the time-step state within HBM, and its restore.
 *   exo_step_redo: nsteps accepted SSPRK3 steps (CFL time step) with that
 *     redo rule; *redos = steps redone by this call.  max_redo = 0 steps as
 *     exo_step does, without checks (the reference's default). */
EXO_API int exo_state_valid(exo_handle *h, int *valid);
EXO_API int exo_save_states(exo_handle *h);
EXO_API int exo_load_states(exo_handle *h);
EXO_API int exo_step_redo(exo_handle *h, int nsteps, int max_redo, double *dt_last, int *redos);

/* --- time-step state ------------------------------------------------------ */
/* dt of the last step and the previous dt that bounds the next CFL time step
 * to twice it (Mesh::NewTimeStep, upstream Athena; patches/21.new_blockdt.patch
 * computes the block limit).  exo_upload_state resets it to (0, unbounded).  A
 * host that restarts from its own restart file (which stores Mesh::dt) sets
 * dt_prev to that dt so the run continues with the same time steps. */
EXO_API int exo_get_dt_state(exo_handle *h, double *dt, double *dt_prev);
EXO_API int exo_set_dt_state(exo_handle *h, double dt, double dt_prev);

/* --- checkpoint / restart ------------------------------------------------- */
/* Replaces the restart round trip of the reference (RestartOutput writing
 * <basename>.rst, and Mesh(pinput, restartfile) reading it in mesh_setup,
 * src/mesh_setup.cpp:31-39,64-69) for this instance: the conserved state
 * with its ghost cells, the passive scalars, the time-step state (dt and the
 * previous dt that bounds its growth), the mixed form's per-level reference,
 * and the caller's time and cycle count.  A run restarted from a checkpoint
 * continues bit for bit as the uninterrupted run.  The file (little-endian) is
 * "EXOCKPT1", a count of metadata doubles (19: the exo_config fields, time,
 * dt, previous dt, cycle-count bits), the payload's three double counts, the
 * payload (u, s and the reference levels in the host layout of
 * exo_upload_state) and an FNV-1a checksum of metadata and payload words.  It
 * is written to <path>.tmp and renamed, so an interrupted save leaves any
 * earlier checkpoint at <path> intact.  Loading requires the same
 * configuration except for the precision (a checkpoint of one precision
 * restarts in the other; an FP64 checkpoint gives a mixed handle the
 * reference levels of the restored state). */
EXO_API int exo_save_checkpoint(exo_handle *h, const char *path, double time, int64_t ncycle);
EXO_API int exo_load_checkpoint(exo_handle *h, const char *path, double *time, int64_t *ncycle);

#ifdef __cplusplus
}
#endif
#endif /* EXOCUBED_H_ */
