/*
 * exo_oracle.h -- CPU FP64 restatement of ExoCubed's cubed-sphere hydro update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * reported CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (exocubed_amd/) never links it.
 *
 * It follows the reference loop by loop, on Athena's per-MeshBlock layout
 * (NVAR, nc3, nc2, nc1) with x1 (radial) fastest, one MeshBlock per panel
 * (mesh level 2).  Each function cites the reference file:line it restates.
 *
 * Parity status per function (see also the header of exo_oracle.c):
 *   pinned by golden vectors : or_weno5 / or_weno3 (tests/test_weno.cpp:11-29)
 *   pinned by closed forms   : geometry (areas/volumes), vector transforms,
 *                              ghost remap of constant / smooth fields,
 *                              block-Thomas vs dense solve (5x5 and 3x3), the
 *                              3x3 partial correction against the decoupled 5x5
 *                              one, Riemann consistency,
 *                              W92 user outputs (known answers, rigid lat-lon
 *                              map, wind speed and direction), the redo path
 *                              (plain stepping without failures; a redone step
 *                              is a step at dt / 2^k from the restored state)
 *   pinned by physics        : (tests/oracle_checks.py, each shown to catch
 *                              planted mistakes by tests/test_oracle_mutants.py)
 *                              the vector remap of a solid-body rotation, the
 *                              L/R exchange, the hydrostatic decomposition,
 *                              Riemann closed forms (Toro ch. 9-10), Coriolis in
 *                              Williamson test 2, the CFL closed form, the
 *                              SSPRK3 composition and finite-volume mass change,
 *                              column conservation of the implicit correction
 *   parity unpinned          : rounding order only, Athena-side semantics not in
 *                              the reference tree (the operation order inside
 *                              AddFluxDivergence and the SSPRK3 average,
 *                              Mesh::NewTimeStep's x2 limit, the reflecting BC
 *                              fill) -- restated from upstream Athena++.
 */
#ifndef EXO_ORACLE_H_
#define EXO_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_EOS_SHALLOW_YZ = 0, OR_EOS_IDEAL = 1 };
enum { OR_RS_ROE_SW = 0, OR_RS_LMARS_SW = 1, OR_RS_HLLC = 2, OR_RS_LMARS = 3 };
enum { OR_COR_NONE = 0, OR_COR_FORCING2 = 1, OR_COR_3D = 2 };

typedef struct or_config {
  int nx;             /* cells per panel edge (N), even */
  int nx1;            /* radial cells */
  int eos;            /* OR_EOS_* */
  int rsolver;        /* OR_RS_* */
  int implicit_flag;  /* bit 1: x1 column solve, bit 8: 5x5 full, else 3x3 partial */
  int ntracer;        /* passive scalars */
  int coriolis;       /* OR_COR_* */
  int cfl_x1;         /* 1: x1 widths enter the CFL minimum */
  double x1min, x1max;
  double gamma;
  double grav;        /* grav_acc1 (negative = downward); 0 for shallow water */
  double omega;       /* planetary rotation rate for the Coriolis source */
  double cfl;
} or_config;

typedef struct or_solver or_solver;

/* scalar kernels (tests/test_weno.cpp) */
double or_weno5(double phim2, double phim1, double phi, double phip1, double phip2);
double or_weno3(double phim1, double phi, double phip1);

/* deterministic transcendental functions shared in algorithm with the device */
double or_exp(double x);
double or_log(double x);
double or_pow(double x, double y);

or_solver *or_create(const or_config *cfg);
void or_destroy(or_solver *s);
int or_nvar(const or_solver *s);
int or_ncells(const or_solver *s, int dim); /* dim 1,2,3 -> nc1,nc2,nc3 */

/* host layout: 6 consecutive MeshBlock arrays, panel p (blockID p+1) first
 * index: [p][n][k][j][i] with (nc3, nc2, nc1) = (N+6, N+6, nx1+6). */
void or_set_cons(or_solver *s, const double *u);
void or_get_cons(const or_solver *s, double *u);
void or_get_prim(const or_solver *s, double *w);
void or_set_scalars(or_solver *s, const double *sc); /* conserved s = rho*r */
void or_get_scalars(const or_solver *s, double *sc);
void or_get_scalar_prim(const or_solver *s, double *r); /* mixing ratios r = s / rho */

/* geometry tables, [p][k][j] of nc3*nc2 (cell) values; which: 0 cos, 1 sin,
 * 2 cos_face2, 3 sin_face2, 4 cos_face3, 5 sin_face3, 6 dx2f_ang, 7 dx3f_ang,
 * 8 dx2f_ang_face3, 9 dx3f_ang_face2 */
void or_get_geom2d(const or_solver *s, int which, double *out);
/* 1-D coordinates: which 0 x1f(nc1+1) 1 x1v(nc1) 2 x2f(nc2+1) 3 x2v 4 x3f 5 x3v
 * for panel p */
void or_get_coord(const or_solver *s, int p, int which, double *out);
/* face areas / volume of one cell of panel p */
double or_cell_volume(const or_solver *s, int p, int k, int j, int i);
double or_face_area(const or_solver *s, int p, int dir, int k, int j, int i);
/* lat/lon of a cell center (radians) */
void or_latlon(const or_solver *s, int p, int k, int j, double *lat, double *lon);
/* user outputs of examples/2023-Chen-exo3/W92.cpp:115-141: lat, lon (degrees),
 * U, V, sqrtg at one point, and at every cell: out (6, 5, nc3, nc2, nc1) */
void or_w92_user_out_point(double x2v, double x3v, int blockID, double Vy, double Vz,
                           double *out5);
void or_user_output(const or_solver *s, double *out);

/* pipeline pieces */
void or_fill_ghosts(or_solver *s);          /* A7 panel remap on u (and s) */
void or_cons_to_prim(or_solver *s);         /* A12 + reflecting x1 BC */
void or_calc_fluxes(or_solver *s);          /* A1+A2+A3+A6+A9 from w */
void or_get_flux(const or_solver *s, int dir, double *out); /* [p][n][..] */
double or_new_dt(or_solver *s);             /* A14 */
void or_stage(or_solver *s, int stage, double dt); /* one SSPRK3 stage */
/* single phases of a stage: -div F + sources times wght with stage-1 weights
 * (into u, or into du for the ideal gas), and the implicit correction */
void or_add_flux_divergence(or_solver *s, double wght);
void or_add_flux_divergence_stage(or_solver *s, int stage, double wght);
void or_implicit_correct(or_solver *s, double dt_stage);
double or_step(or_solver *s, int nsteps, double dt_fixed); /* returns last dt */
/* the main loop's redo path (main.cpp.in:126-160): nsteps accepted steps,
 * an invalid step undone and repeated with half the dt; returns the steps
 * redone, -1 after too many */
int or_step_redo(or_solver *s, int nsteps, int max_redo, double *dt_last);
int or_state_valid(const or_solver *s); /* last primitive conversion valid */
void or_set_threads(int n);

/* single-column / single-row shims used by the per-function parity tests */
void or_implicit_column(const or_config *cfg, int nx1, const double *w_col,
                        /* (5, nx1+6) prim of one column incl. x1 ghosts */
                        double *du_col, /* (5, nx1+6) in/out, interior used */
                        const double *area1, /* nx1+1 face areas */
                        const double *vol,   /* nx1 volumes */
                        double dt);
void or_riemann_face(const or_config *cfg, int ivx, const double *wl,
                     const double *wr, double *flx, double cth_face,
                     double cos_face, double sin_face, double cos_cell);

/* the hydrostatic decomposition of one column of the current state
 * (change_to_perturbation.cpp:39-190): perturbations, face pressures, the
 * restored column and the restored face states of zero perturbations */
void or_decompose_column(const or_solver *s, int p, int k, int j, double *wpert, double *psf,
                         double *wrest, double *face_p, double *face_rho);
/* pins of the implicit-solve algebra and the vector maps */
void or_flux_jacobian(double gm1, const double *w, double *out25);
void or_eigensystem(const double *prim, double cs, double gm1, double *R25, double *Ri25,
                    double *absA25);
void or_roe_average(const double *wl, const double *wr, double gm1, double *prim);
/* the partial (3x3: IDN, IVX, IEN) column correction and its block-Thomas */
void or_implicit_column3(const or_config *cfg, int nx1, const double *w_col, double *du_col,
                         const double *area1, const double *vol, double dt);
void or_block_thomas3(int n, const double *a, const double *b, const double *c, const double *d,
                      double *x);
void or_block_thomas(int n, const double *a, const double *b, const double *c,
                     const double *rhs, double *x);
void or_vec_maps(double X, double Y, int blockID, double V2, double V3, double *U, double *V,
                 double *V2b, double *V3b);

#ifdef __cplusplus
}
#endif
#endif /* EXO_ORACLE_H_ */
