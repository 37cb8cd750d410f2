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

#define EXO_OK 0
#define EXO_ERR_ARG (-1)
#define EXO_ERR_HIP (-2)
#define EXO_ERR_STATE (-3)
#define EXO_ERR_NOMEM (-4)
#define EXO_ERR_UNSUPPORTED (-5)

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
  int implicit_flag; /* 0 or 9 (<hydro> implicit_flag, solve_implicit_3d.cpp:102-135) */
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
int exo_create(const exo_config *cfg, exo_handle **out);
int exo_destroy(exo_handle *h);
const char *exo_last_error(const exo_handle *h);
/* number of conserved variables (NHYDRO) and padded cells per panel */
int exo_nvar(const exo_handle *h);
size_t exo_state_doubles(const exo_handle *h); /* 6*NVAR*nc3*nc2*nc1 */

/* --- state transfer (host layout above) -------------------------------- */
/* Upload conserved variables u and (if ntracer > 0) conserved scalars s = rho r
 * (NTRACER, nc3, nc2, nc1 per panel; may be NULL).  Interior cells are read;
 * ghosts are then filled as Mesh::Initialize does on the cube (ghost exchange,
 * patches/29.mesh.patch:19-86, then cons->prim and the reflecting x1 BC). */
int exo_upload_state(exo_handle *h, const double *u, const double *s);
int exo_download_state(exo_handle *h, double *u, double *s);
int exo_download_prim(exo_handle *h, double *w);

/* --- the hot path, one call per reference entry point ------------------- */
/* Hydro::CalculateFluxes (patches/19.calculate_fluxes.patch:17-163): WENO5
 * reconstruction (Reconstruction::Weno5X1/X2/X3), hydrostatic decomposition
 * around the x1 sweep (change_to_perturbation.cpp:39-190), panel-edge L/R
 * exchange (cubed_sphere.cpp:267-687) and Hydro::RiemannSolver on every face. */
int exo_calc_fluxes(exo_handle *h);
/* Copy the face fluxes of direction dir (0 = x1, 1 = x2, 2 = x3) to the host in
 * Athena's flux layout: x1 (NVAR, nc3, nc2, nc1+1), x2 (NVAR, nc3, nc2+1, nc1),
 * x3 (NVAR, nc3+1, nc2, nc1) per panel; only interior faces are defined. */
int exo_download_flux(exo_handle *h, int dir, double *out);
/* Hydro::AddFluxDivergence(wght, u) followed by
 * GnomonicEquiangle::AddCoordTermsDivergence (gnomonic_equiangle.cpp:834-879)
 * into u (shallow water) or into the tendency du (ideal gas, zeroed first, as
 * ImplicitHydroTasks::IntegrateHydro, implicit_hydro_tasks.cpp:157-162). */
int exo_add_flux_divergence(exo_handle *h, double wght);
/* CubedSphereUtility::PackData panel-edge remap (cubed_sphere_utility.cpp:
 * 371-617) of the conserved state into the face ghost zones, then
 * cons->prim (EquationOfState::ConservedToPrimitive) and the reflecting x1 BC. */
int exo_fill_ghosts(exo_handle *h);
/* ImplicitSolver::SolveImplicit3D (solve_implicit_3d.cpp:27-180) with
 * FullCorrection on the x1 columns, applied to du, then u += du
 * (ImplicitHydroTasks::ImplicitCorrection, implicit_hydro_tasks.cpp:245-266). */
int exo_implicit_correct(exo_handle *h, double dt_stage);
/* Hydro::NewBlockTimeStep (patches/21.new_blockdt.patch:16-37) and
 * Mesh::NewTimeStep's factor-2 growth limit; returns the next dt. */
int exo_new_dt(exo_handle *h, double *dt);
/* One SSPRK3 stage (1..3) with time step dt, as the task list runs it. */
int exo_stage(exo_handle *h, int stage, double dt);
/* nsteps full SSPRK3 steps, dt from the CFL condition on the device when
 * dt_fixed <= 0; *dt_last (may be NULL) receives the last step's dt.
 * Replaces the main loop body (main.cpp.in:130-186). */
int exo_step(exo_handle *h, int nsteps, double dt_fixed, double *dt_last);
/* Block until all work queued on the handle's stream has finished. */
int exo_sync(exo_handle *h);
/* Live kernel timing: between begin and end every launch on the handle's stream
 * is bracketed by a HIP event pair; end returns the summed milliseconds and
 * launch counts per kernel id (0 prim, 1 ghost, 2 flux_x1, 3 edge_states,
 * 4 flux_x2, 5 flux_x3, 6 update, 7 implicit, 8 dt). */
int exo_profile_begin(exo_handle *h);
int exo_profile_end(exo_handle *h, double *ms, int *counts, int nkinds);

/* --- problem set-up (synthetic initial states of the benchmark configs) - */
/* Fill host conserved arrays (interior cells) for a seeded initial state.
 * kind 0: shallow-water random-mode height field in geostrophic balance
 *         (params: mean depth, mode amplitude, g);
 * kind 1: isothermal hydrostatic atmosphere with Gaussian temperature noise
 *         and random-phase zonal waves (params: T0, ps, Rd, noise fraction,
 *         wave amplitude);
 * kind 2: lapse-rate atmosphere to an isothermal stratosphere (params: Ts,
 *         lapse K/m, Tstrat, ps, Rd, noise K). */
int exo_init_state(exo_handle *h, int kind, uint64_t seed, const double *params,
                   int nparams, double *u_out, double *s_out);
/* The same without a handle (host only, no GPU needed). */
int exo_init_state_cfg(const exo_config *cfg, int kind, uint64_t seed, const double *params,
                       int nparams, double *u_out, double *s_out);

/* --- single-face and single-column shims for per-function parity ------- */
/* Riemann flux of one face with the given L/R primitive states, computed by
 * the device kernel (dir 0/1/2 = x1/x2/x3; the metric of face (p,k,j) of this
 * handle's grid is used). */
int exo_riemann_face(exo_handle *h, int dir, int p, int k, int j, const double *wl,
                     const double *wr, double *flx);
/* WENO5 value at the face between b and c seen from cell c (device). */
int exo_weno5_batch(exo_handle *h, const double *in5, double *out, int n);
/* deterministic device exp/log/pow on n inputs (op 0 exp, 1 log, 2 pow y) */
int exo_math_batch(exo_handle *h, int op, const double *x, double y, double *out, int n);

/* Host-only introspection of the set-up tables (no GPU needed): which 0 ghost
 * remap descriptors, 1 edge-exchange descriptors, 2 metric tables
 * [16][6][nc3][nc2], 3 x1f|x1v|dx1f, 4 x2f and 5 x3f of the six panels.
 * With buf == NULL only *count is returned. */
int exo_describe(const exo_config *cfg, int which, void *buf, size_t cap_bytes,
                 size_t *count);

#ifdef __cplusplus
}
#endif
#endif /* EXOCUBED_H_ */
