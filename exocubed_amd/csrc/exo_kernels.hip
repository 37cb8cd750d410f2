// exo_kernels.hip -- gfx950 kernels of the cubed-sphere finite-volume update
// and the C ABI of include/exocubed.h.
//
// Device layout of every cell-centred array of one instance:
//   [var][panel][k][i][j], j (x2) fastest, row pitch jp >= nc2 + 1
// so that wavefront lanes walk j: x1 column sweeps and x3 sweeps stream
// coalesced rows, x2 sweeps read neighbouring lanes' rows.  Face fluxes use
// the same layout with one extra k plane and one extra i row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/exocubed.h"
#include "exo_device.hpp"
#include "exo_mixed.hpp"
#include "exo_netcdf.hpp"
#include "exo_setup.hpp"

using namespace exo;

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) throw std::string("HIP: ") + hipGetErrorString(e_) + \
                                " at " #x;                                     \
  } while (0)

namespace {

// ---------------------------------------------------------------------------
// Kernel-side view of one instance.  AS is the address space of the array
// pointers: 0 (generic) where the host fills the structures, 1 (global) in the
// copies the kernels work on, so that every array access is a global (not a
// flat) memory instruction.
#define ADDR(a) __attribute__((address_space(a)))

// pointer to T in address space AS (0: the language's default pointer)
template <int AS, typename T>
struct APtr {
  using type = ADDR(AS) T *;
};
template <typename T>
struct APtr<0, T> {
  using type = T *;
};
#define PTR(T) typename APtr<AS, T>::type

template <int AS>
struct GridT {
  int N, nx1, nc1, nc2, nc3, jp, nv, nt;
  int is, ie, js, je, ks, ke;
  int eos, rs, coriolis, implicit, cfl_x1;
  double gamma, grav, cfl;
  size_t plane, pstride, vstride;      // cell arrays
  size_t fplane, fpstride, fvstride;   // face arrays
  // geometry
  PTR(const double) x1f, x1v, dx1f;
  PTR(const double) tab;               // [T_NTAB][6][nc3][nc2]
  PTR(const double) tab1;              // [T1_NTAB][6][nc2 + 1] per-row tangent factors
  size_t t2stride;                     // 6*nc3*nc2
  __host__ __device__ size_t c(int v, int p, int k, int i, int j) const {
    return v * vstride + p * pstride + k * plane + (size_t)i * jp + j;
  }
  // variables a kernel can skip: the shallow-water x1 velocity is zero for
  // good (shallow_yz_hydro.cpp:57 sets it, every flux of it is zero), so it is
  // never reconstructed, fluxed, updated or stored after the first fill
  __host__ __device__ bool skipv(int n) const { return n >= nv || (eos == 0 && n == 1); }
  __host__ __device__ size_t f(int v, int p, int k, int i, int j) const {
    return v * fvstride + p * fpstride + k * fplane + (size_t)i * jp + j;
  }
  __device__ double T(int t, int p, int k, int j) const {
    return tab[t * t2stride + ((size_t)p * nc3 + k) * nc2 + j];
  }
  // row q (j for the x2 factors, k for the x3 factors; the panel is square)
  __device__ double T1(int t, int p, int q) const {
    return tab1[((size_t)t * 6 + p) * (nc2 + 1) + q];
  }
  __device__ double area1(int p, int k, int j, int i) const {
    return T(T_DX2A, p, k, j) * T(T_DX3A, p, k, j) * x1f[i] * x1f[i] * T(T_SINC, p, k, j);
  }
  __device__ double area2(int p, int k, int j, int i) const {
    return x1v[i] * T(T_DX3A_F2, p, k, j) * dx1f[i];
  }
  __device__ double area3(int p, int k, int j, int i) const {
    return x1v[i] * T(T_DX2A_F3, p, k, j) * dx1f[i];
  }
  __device__ double volume(int p, int k, int j, int i) const {
    return dx1f[i] * T(T_DX2A, p, k, j) * T(T_DX3A, p, k, j) * x1v[i] * x1v[i] *
           T(T_SINC, p, k, j);
  }
};

template <int AS>
struct StateT {
  PTR(double) u, w, u1, du;             // nv cell arrays
  PTR(double) s, r, s1;                 // nt cell arrays
  PTR(double) flx[3], sflx[3];          // face arrays
  PTR(double) edge;                     // [nv][6][4][N][nx1] ghost-side face states
  PTR(double) coef;                     // implicit column coefficients, 30 per cell
  PTR(unsigned long long) dtbits;       // CFL minimum as ordered bits
  PTR(double) dtv;                      // [0] dt of the step, [1] previous dt
  // mixed precision: FP32 perturbation fluxes, FP64 reference parts
  PTR(float) fflx[3], fsflx[3];         // face fluxes without the reference pressure
  PTR(float) fedge;                     // ghost-side edge states (perturbations)
  PTR(double) psf;                      // FP64 hydrostatic x1 face pressure [p][k][i][j]
  PTR(double) lref;                     // per level: [0,nc1) split pressure P, [nc1,2nc1)
                                        // the reconstructed variable's reference,
                                        // [2nc1,3nc1) rho_ref
  PTR(double) part;                     // per-block compensated partial sums [block][5][2]
  PTR(float) fw;                        // mixed ideal gas: FP32 perturbation primitives
                                        // (w - level reference) and tracer r, [nv + nt]
                                        // cell arrays, written with w by k_prim
  PTR(int) bad;                         // set by the primitive conversion of the stage
                                        // when a density or pressure is negative or NaN
};

using Grid = GridT<0>;
using State = StateT<0>;
using DGrid = GridT<1>;
using DState = StateT<1>;

template <typename T>
__device__ __forceinline__ ADDR(1) T *glb(T *p) {
  return (ADDR(1) T *)p;
}

__device__ __forceinline__ DGrid to_global(const Grid &h) {
  DGrid g;
  g.N = h.N; g.nx1 = h.nx1; g.nc1 = h.nc1; g.nc2 = h.nc2; g.nc3 = h.nc3; g.jp = h.jp;
  g.nv = h.nv; g.nt = h.nt;
  g.is = h.is; g.ie = h.ie; g.js = h.js; g.je = h.je; g.ks = h.ks; g.ke = h.ke;
  g.eos = h.eos; g.rs = h.rs; g.coriolis = h.coriolis; g.implicit = h.implicit;
  g.cfl_x1 = h.cfl_x1;
  g.gamma = h.gamma; g.grav = h.grav; g.cfl = h.cfl;
  g.plane = h.plane; g.pstride = h.pstride; g.vstride = h.vstride;
  g.fplane = h.fplane; g.fpstride = h.fpstride; g.fvstride = h.fvstride;
  g.x1f = glb(h.x1f); g.x1v = glb(h.x1v); g.dx1f = glb(h.dx1f); g.tab = glb(h.tab);
  g.tab1 = glb(h.tab1);
  g.t2stride = h.t2stride;
  return g;
}

__device__ __forceinline__ DState to_global(const State &h) {
  DState s;
  s.u = glb(h.u); s.w = glb(h.w); s.u1 = glb(h.u1); s.du = glb(h.du);
  s.s = glb(h.s); s.r = glb(h.r); s.s1 = glb(h.s1);
  for (int d = 0; d < 3; ++d) {
    s.flx[d] = glb(h.flx[d]);
    s.sflx[d] = glb(h.sflx[d]);
    s.fflx[d] = glb(h.fflx[d]);
    s.fsflx[d] = glb(h.fsflx[d]);
  }
  s.edge = glb(h.edge); s.coef = glb(h.coef); s.dtbits = glb(h.dtbits); s.dtv = glb(h.dtv);
  s.fedge = glb(h.fedge); s.psf = glb(h.psf); s.lref = glb(h.lref); s.part = glb(h.part);
  s.bad = glb(h.bad);
  s.fw = glb(h.fw);
  return s;
}

constexpr int BLK = 256;

// One instance as the kernels see it, and the work list of one launch.  The
// blocks of a launch are split over the instances by a prefix table (an
// instance that a kernel does not apply to gets no blocks); a launch with
// fewer workgroups than blocks loops over the rest (a persistent launch), so
// many small instances of different sizes run in one launch per phase.
struct Inst {
  Grid g;
  State st;
  const GhostDesc *gd;
  const EdgeDesc *ed;
  int ngd, ned;
};

// The instance list itself is a separate __restrict__ kernel argument: no
// store of a kernel can then alias it, so the pointers read from it are known
// to be global memory and are accessed with global (not flat) instructions.
struct Work {
  const int *bstart;   // [ninst + 1] first block of each instance
  int ninst, nblk;
};

// the instance owning block b: the last one whose first block is <= b
__device__ __forceinline__ int find_inst(const Work &w, int b) {
  int lo = 0, hi = w.ninst - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (w.bstart[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup h runs on
// XCD h mod 8), each with its own L2.  xcd_order gives XCD x the contiguous
// range [x q + min(x, r), ...) of the n blocks (q = n / 8, r = n mod 8), so
// neighbouring rows and planes of a stencil are read through one L2.  It is a
// bijection of [0, n).
__device__ __forceinline__ int xcd_order(int h, int n) {
  int q = n >> 3, r = n & 7, x = h & 7;
  return x * q + (x < r ? x : r) + (h >> 3);
}

// for each block of the work list this workgroup owns: g, st and the block's
// index lb within its instance
#define WORK_LOOP(w)                                                     \
  for (int h_ = blockIdx.x; h_ < (w).nblk; h_ += gridDim.x)
#define WORK_INST(w)                                                     \
  int b_ = xcd_order(h_, (w).nblk);                                      \
  int ii_ = find_inst((w), b_);                                          \
  const DGrid g = to_global(I_[ii_].g);                                  \
  const DState st = to_global(I_[ii_].st);                               \
  int lb = b_ - (w).bstart[ii_]

// ---------------------------------------------------------------------------
// EquationOfState::ConservedToPrimitive (shallow_yz_hydro.cpp:42-67,
// ideal_moist_hydro.cpp:77-159) on the interior levels of interior and
// face-ghost columns, then the reflecting x1 boundary (v1 mirrored): one
// thread per cell; a cell within NG levels of either end also writes the
// ghost it mirrors into (ghost is - q mirrors is + q - 1, ghost ie + q mirrors
// ie - q + 1).  With nx1 < NG the ghosts beyond the mirror of an interior cell
// are not written (only shallow water allows it, and it never reads x1 ghosts).
// The equation of state of a kernel specialised for one (EK = EOS_SW or
// EOS_IDEAL: the compiler drops the other's code and registers) or read from
// the instance (EK < 0), and whether it skips variable n.
template <int EK>
__device__ __forceinline__ int eos_of(const DGrid &g) {
  return EK >= 0 ? EK : g.eos;
}
template <int EK>
__device__ __forceinline__ bool skip_of(const DGrid &g, int n) {
  return EK == EOS_SW ? n == IVX || n > IVZ : EK == EOS_IDEAL ? false : g.skipv(n);
}
template <int EK>
__device__ __forceinline__ int nv_of(const DGrid &g) {
  return EK == EOS_SW ? 4 : EK == EOS_IDEAL ? 5 : g.nv;
}
// The Riemann solver of a kernel specialised for one (RK = RS_*) or read from
// the instance (RK < 0); a fixed solver fixes the equation of state too (the
// configuration pairs RS_ROE_SW and RS_LMARS_SW with shallow water).
template <int RK>
__device__ __forceinline__ int rs_of(const DGrid &g) {
  return RK >= 0 ? RK : g.rs;
}
template <int EK, int RK>
constexpr int eos_k() {
  return RK < 0 ? EK : RK <= RS_LMARS_SW ? EOS_SW : EOS_IDEAL;
}

// Primitive state of one cell (p, k, i, j) from its conserved values uv
// (EquationOfState::ConservedToPrimitive of shallow_yz_hydro.cpp /
// ideal_moist_hydro.cpp): w, the reflecting x1 mirrors and the scalar mixing
// ratios; the shallow-water depth floor is written back to u.
template <int EK = -1>
__device__ __forceinline__ void prim_from(const DGrid &g, const DState &st, int p, int k, int i,
                                          int j, const double *uv, double dfloor, double sfloor,
                                          double cth) {
  double wv[5] = {0., 0., 0., 0., 0.};
  double rho_c;
  if (eos_of<EK>(g) == EOS_SW) {
    double ud = uv[IDN];
    if (!(ud > dfloor)) {  // the floor, written back only where it applies
      ud = dfloor;
      st.u[g.c(IDN, p, k, i, j)] = ud;
    }
    double v = uv[IVY], ww = uv[IVZ];
    double di = qdiv(1., ud);
    // the four quotients by sth2 share one correctly rounded reciprocal
    double sth2 = 1. - cth * cth, ys = qdiv(1., sth2);
    double wy = div_rn(v, sth2, ys) - div_rn(ww * cth, sth2, ys);
    double wz = div_rn(-v * cth, sth2, ys) + div_rn(ww, sth2, ys);
    rho_c = ud;
    wv[IDN] = ud;
    wv[IVX] = 0.;
    wv[IVY] = wy * di;
    wv[IVZ] = wz * di;
  } else {
    double gm1 = g.gamma - 1.0;
    double u_d = uv[IDN], u_m1 = uv[IVX];
    double u_m2 = uv[IVY], u_m3 = uv[IVZ];
    double u_e = uv[IEN];
    double density = 0.;
    density += u_d;
    rho_c = u_d;
    double di = qdiv(1., density);
    double w_vx = u_m1 * di, w_vy = u_m2 * di, w_vz = u_m3 * di;
    double sth2 = 1. - cth * cth, ys = qdiv(1., sth2);
    double v = w_vy, ww = w_vz;
    w_vy = div_rn(v, sth2, ys) - div_rn(ww * cth, sth2, ys);
    w_vz = div_rn(-v * cth, sth2, ys) + div_rn(ww, sth2, ys);
    double KE = 0.5 * (u_m1 * w_vx + u_m2 * w_vy + u_m3 * w_vz);
    wv[IDN] = density;
    wv[IVX] = w_vx;
    wv[IVY] = w_vy;
    wv[IVZ] = w_vz;
    wv[IPR] = gm1 * (u_e - KE) * 1. / 1.;
    // ideal_moist_hydro.cpp:108,150: a negative or NaN density or pressure
    // marks the state invalid (the shallow-water EOS has no such check)
    if (!(density >= 0.) || !(wv[IPR] >= 0.)) *st.bad = 1;
  }
  // mirrors into ghosts is - qb, ie + qt (shallow water never reads x1 ghosts)
  int qb = i - g.is + 1, qt = g.ie - i + 1;
  bool mirror = eos_of<EK>(g) != EOS_SW;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (skip_of<EK>(g, n)) continue;
    st.w[g.c(n, p, k, i, j)] = wv[n];
    double m = n == IVX ? -wv[n] : wv[n];
    if (mirror && qb <= NG) st.w[g.c(n, p, k, g.is - qb, j)] = m;
    if (mirror && qt <= NG) st.w[g.c(n, p, k, g.ie + qt, j)] = m;
  }
  // the mixed ideal gas's face and edge kernels read these values as FP32
  // perturbations about the level reference of pressure and density
  if (st.fw) {
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      double v = wv[n];
      if (n == IPR) v -= st.lref[g.nc1 + i];
      else if (n == IDN) v -= st.lref[2 * g.nc1 + i];
      st.fw[g.c(n, p, k, i, j)] = (float)v;
    }
  }
  for (int n = 0; n < g.nt; ++n) {
    double rr = st.s[g.c(n, p, k, i, j)] * qdiv(1.0, rho_c);
    rr = rr < sfloor ? sfloor : rr;
    st.r[g.c(n, p, k, i, j)] = rr;
    if (qb <= NG) st.r[g.c(n, p, k, g.is - qb, j)] = rr;
    if (qt <= NG) st.r[g.c(n, p, k, g.ie + qt, j)] = rr;
    if (st.fw) st.fw[g.c(g.nv + n, p, k, i, j)] = (float)rr;
  }
}


// u is loaded non-temporally: nothing reads it again before the stage's
// column step, and the caches are left to w, which the flux kernels and the
// CFL minimum read next (k_dt 0.055 -> 0.044 ms on 6 x 96 x 96 x 64)
template <int EK = -1>
__device__ __forceinline__ void prim_body(const DGrid &g, const DState &st, int lb, double dfloor,
                                          double sfloor) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int ni = g.nx1;
  size_t total = (size_t)6 * g.nc3 * ni * g.nc2;
  if (t >= total) return;
  int j = (int)(t % g.nc2);
  size_t rr0 = t / g.nc2;
  int i = g.is + (int)(rr0 % ni);
  rr0 /= ni;
  int k = (int)(rr0 % g.nc3), p = (int)(rr0 / g.nc3);
  bool kin = k >= g.ks && k <= g.ke, jin = j >= g.js && j <= g.je;
  if (!kin && !jin) return;  // corner ghosts are never read
  double uv[5];
#pragma unroll
  for (int n = 0; n < 5; ++n)
    uv[n] = skip_of<EK>(g, n) ? 0. : __builtin_nontemporal_load(&st.u[g.c(n, p, k, i, j)]);
  prim_from<EK>(g, st, p, k, i, j, uv, dfloor, sfloor, g.T(T_COSC, p, k, j));
}

// The ghost strips only (3 deep along each panel edge, corners excluded): the
// primitive state of the interior cells of shallow water is written by the
// stage's update (shallow-water instances only).  Threads cover
// [p][side][depth][along][i].
__device__ __forceinline__ void prim_ghost_body(const DGrid &g, const DState &st, int lb,
                                                double dfloor, double sfloor) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * 4 * NG * N * ni;
  if (t >= total) return;
  int a = (int)(t % N);
  size_t r = t / N;
  int i = g.is + (int)(r % ni);
  r /= ni;
  int dq = (int)(r % NG);
  r /= NG;
  int side = (int)(r % 4), p = (int)(r / 4);
  int k, j;
  if (side == 0) { j = g.js - 1 - dq; k = g.ks + a; }
  else if (side == 1) { j = g.je + 1 + dq; k = g.ks + a; }
  else if (side == 2) { k = g.ks - 1 - dq; j = g.js + a; }
  else { k = g.ke + 1 + dq; j = g.js + a; }
  // plain loads: the ghost remap has just written these cells
  double uv[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) uv[n] = skip_of<EOS_SW>(g, n) ? 0. : st.u[g.c(n, p, k, i, j)];
  prim_from<EOS_SW>(g, st, p, k, i, j, uv, dfloor, sfloor, g.T(T_COSC, p, k, j));
}

// EK: the equation of state of every instance of the launch, or -1 (any)
template <int EK = -1>
__global__ void __launch_bounds__(BLK) k_prim(const Inst *__restrict__ I_, Work w, double dfloor,
                                              double sfloor) {
  WORK_LOOP(w) {
    WORK_INST(w);
    prim_body<EK>(g, st, lb, dfloor, sfloor);
  }
}

__global__ void __launch_bounds__(BLK) k_prim_ghost(const Inst *__restrict__ I_, Work w, double dfloor,
                                                    double sfloor) {
  WORK_LOOP(w) {
    WORK_INST(w);
    prim_ghost_body(g, st, lb, dfloor, sfloor);
  }
}

// ---------------------------------------------------------------------------
// Panel-edge ghost remap (CubedSphereUtility::PackData with InteprolateX2/X3,
// cubed_sphere_utility.cpp:92-617): one thread per (ghost cell, i); the
// descriptor carries the interpolation weights and vector-map points.
__device__ __forceinline__ void ghost_body(const DGrid &g, int lb,
                                           const ADDR(1) GhostDesc *__restrict__ gd, int ndesc,
                                           ADDR(1) double *arr, int nvar, int vectors) {
  int t = lb * BLK + threadIdx.x;
  int ni = g.nx1;
  if (t >= ndesc * ni) return;
  int di = t / ni, i = g.is + t % ni;
  GhostDesc d = *(const GhostDesc *)(gd + di);
  int sp = d.src_panel;
  size_t o0, o1;
  if (d.along_x2) {
    o0 = g.c(0, sp, d.src_row, i, d.src_a0);
    o1 = g.c(0, sp, d.src_row, i, d.src_a0 + 1);
  } else {
    o0 = g.c(0, sp, d.src_a0, i, d.src_row);
    o1 = g.c(0, sp, d.src_a0 + 1, i, d.src_row);
  }
  size_t ot = g.c(0, d.tgt_panel, d.tgt_k, i, d.tgt_j);
  for (int n = 0; n < nvar; ++n) {
    if (vectors && (n == IVY || n == IVZ || g.skipv(n))) continue;
    double v1 = arr[n * g.vstride + o0], v2 = arr[n * g.vstride + o1];
    arr[n * g.vstride + ot] = qdiv(d.A * v1 + d.B * v2, d.D);
  }
  if (vectors) {
    double v1y = arr[IVY * g.vstride + o0], v1z = arr[IVZ * g.vstride + o0];
    double v2y = arr[IVY * g.vstride + o1], v2z = arr[IVZ * g.vstride + o1];
    double vy = qdiv(d.A * v1y + d.B * v2y, d.D);
    double vz = qdiv(d.A * v1z + d.B * v2z, d.D);
    double sth = d.sth_s, cth = d.cth_s;
    double vyt = vy, vzt = vz;
    vy = qdiv(qdiv(vyt, sth), sth) - qdiv(qdiv(vzt * cth, sth), sth);
    vz = qdiv(qdiv(vzt, sth), sth) - qdiv(qdiv(vyt * cth, sth), sth);
    double U, V;
    vec_rll_from_abp(d.src_cy, d.src_cz, d.src_block, vy, vz, &U, &V);
    vec_abp_from_rll(d.tgt_cy, d.tgt_cz, d.tgt_block, U, V, &vy, &vz);
    sth = d.sth_t;
    cth = d.cth_t;
    vyt = vy;
    vzt = vz;
    vy = vyt + vzt * cth;
    vz = vzt + vyt * cth;
    arr[IVY * g.vstride + ot] = vy;
    arr[IVZ * g.vstride + ot] = vz;
  }
}

// which = 0: conserved variables (vectors remapped), 1: passive scalars
__global__ void __launch_bounds__(BLK) k_ghost(const Inst *__restrict__ I_, Work w, int which) {
  WORK_LOOP(w) {
    WORK_INST(w);
    const Inst &I = I_[ii_];
    if (which == 0) ghost_body(g, lb, glb(I.gd), I.ngd, st.u, g.nv, 1);
    else ghost_body(g, lb, glb(I.gd), I.ngd, st.s, g.nt, 0);
  }
}

// ---------------------------------------------------------------------------
// x1 (radial) fluxes, one thread per interior column: the hydrostatic
// decomposition of Decomposition::ChangeToPerturbation / RestoreFromPerturbation
// (change_to_perturbation.cpp:39-190) around Weno5X1, then the x1 Riemann
// solve (patches/19.calculate_fluxes.patch:77-106).  The face pressure psf
// lives in LDS, [face][lane].
// The block holds 64 columns and X1W waves: wave 0 integrates the face
// pressure of every column into LDS, then each wave takes a contiguous share
// of the faces with its own WENO window, so a column's faces are worked on by
// X1W lanes at once.  Per-face arithmetic is independent of the split.
constexpr int X1W = 4;  // the default; the autotuner may pick 2 or 8 (x1_waves)

__device__ __forceinline__ void psf_scan(const DGrid &g, const DState &st, int p, int k, int j,
                                         double *psf, int lane) {
#define PSF(ii) psf[(ii) * 64 + lane]
  int is = g.is, ie = g.ie;
  double grav = -g.grav;
  double pie = st.w[g.c(IPR, p, k, ie, j)], die = st.w[g.c(IDN, p, k, ie, j)];
  double RdTv = qdiv(pie, die);
  PSF(ie + 1) = pie * dexp(qdiv(-grav * g.dx1f[ie], 2 * RdTv));
  // the densities of eight levels are loaded before the recurrence uses them
  constexpr int SB = 8;
  for (int i0 = ie; i0 >= is; i0 -= SB) {
    double rho[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u)
      rho[u] = i0 - u >= is ? st.w[g.c(IDN, p, k, i0 - u, j)] : 0.;
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      int i = i0 - u;
      if (i >= is) PSF(i) = PSF(i + 1) + grav * rho[u] * g.dx1f[i];
    }
  }
  for (int q = 1; q <= NG; ++q) PSF(is - q) = PSF(is + q);
  for (int q = 1; q <= NG; ++q) PSF(ie + q + 1) = PSF(ie + 1 - q);
#undef PSF
}

template <int W, int RK = -1>
__device__ __forceinline__ void flux_x1_body(const DGrid &g, const DState &st, int lb, double *lds) {
  int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  int N = g.N;
  int nbx = (N * N + 63) / 64;
  int col = (lb % nbx) * 64 + lane;
  int p = lb / nbx;
  bool act = col < N * N;
  int k = g.ks + (act ? col / N : 0), j = g.js + (act ? col % N : 0);
  double *psf = lds;  // (nc1 + 1) x 64
#define PSF(ii) psf[(ii) * 64 + lane]
  int is = g.is, ie = g.ie, nv = g.nv;
  bool decomp = g.grav != 0.;
  double e1 = 1., ig = 1.;
  if (seg == 0 && act && decomp) psf_scan(g, st, p, k, j, psf, lane);
  __syncthreads();
  if (!act) return;
  if (decomp) {
    double ent0 = g.gamma;
    double ent1 = dlog(st.w[g.c(IPR, p, k, is, j)]) - ent0 * dlog(st.w[g.c(IDN, p, k, is, j)]);
    ig = qdiv(1., ent0);
    e1 = dexp(qdiv(-ent1, ent0));
  }
  // this wave's faces [f0, f1)
  int nf = ie + 2 - is;
  int f0 = is + (nf * seg) / W, f1 = is + (nf * (seg + 1)) / W;
  // window of 6 cells (f-3 .. f+2) for face f, perturbed IDN / IPR
  double q[6][5];
  auto load = [&](int slot, int i) {
#pragma unroll
    for (int n = 0; n < 5; ++n) q[slot][n] = n < nv ? st.w[g.c(n, p, k, i, j)] : 0.;
    if (decomp) {
      double a = PSF(i), b = PSF(i + 1), psv;
      if (fabs(a - b) < 1.E-6) psv = (a + b) / 2.;
      else psv = qdiv(a - b, dlog(qdiv(a, b)));
      double dsv = dpow(psv, ig) * e1;
      q[slot][IPR] -= psv;
      q[slot][IDN] -= dsv;
    }
  };
#pragma unroll
  for (int s = 0; s < 6; ++s) load(s, f0 - 3 + s);
  double cth = g.T(T_CTH1, p, k, j), cosc = g.T(T_COSC, p, k, j);
  for (int fi = f0; fi < f1; ++fi) {
    double wl[5], wr[5], fl[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      if (n < nv) {
        wl[n] = weno5(q[4][n], q[3][n], q[2][n], q[1][n], q[0][n]);
        wr[n] = weno5(q[1][n], q[2][n], q[3][n], q[4][n], q[5][n]);
      }
    }
    if (decomp) {
      double pf = PSF(fi);
      wr[IPR] += pf;
      if (wr[IPR] < 0.) wr[IPR] = pf;
      wl[IPR] += pf;
      if (wl[IPR] < 0.) wl[IPR] = pf;
      double dsf = dpow(pf, ig) * e1;
      wr[IDN] += dsf;
      if (wr[IDN] < 0.) wr[IDN] = dsf;
      wl[IDN] += dsf;
      if (wl[IDN] < 0.) wl[IDN] = dsf;
    }
    riemann<double>(EOS_IDEAL, rs_of<RK>(g), IVX, g.gamma, wl, wr, fl, cth, 0., 1., cosc);
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (n < nv) st.flx[0][g.f(n, p, k, fi, j)] = fl[n];
    if (fi + 1 < f1) {
#pragma unroll
      for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int n = 0; n < 5; ++n) q[s][n] = q[s + 1][n];
      load(5, fi + 3);
    }
  }
  // passive scalars: WENO of r along x1, upwinded by the mass flux; a window
  // of 6 levels (fi - 3 .. fi + 2) slides up the column, the next level and
  // mass flux loaded before the face's arithmetic
  for (int n = 0; n < g.nt; ++n) {
    double r[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) r[s] = st.r[g.c(n, p, k, f0 - 3 + s, j)];
    double m = f0 < f1 ? st.flx[0][g.f(IDN, p, k, f0, j)] : 0.;
    for (int fi = f0; fi < f1; ++fi) {
      bool more = fi + 1 < f1;
      double rn = more ? st.r[g.c(n, p, k, fi + 3, j)] : 0.;
      double mn = more ? st.flx[0][g.f(IDN, p, k, fi + 1, j)] : 0.;
      double rl = weno5(r[4], r[3], r[2], r[1], r[0]);
      double rr = weno5(r[1], r[2], r[3], r[4], r[5]);
      st.sflx[0][g.f(n, p, k, fi, j)] = m >= 0.0 ? m * rl : m * rr;
#pragma unroll
      for (int s = 0; s < 5; ++s) r[s] = r[s + 1];
      r[5] = rn;
      m = mn;
    }
  }
#undef PSF
}

// three wavefronts per SIMD (168 VGPRs, no spills): with 4 wavefronts per
// 64-column block the LDS of the face pressures allows four blocks per CU, and
// the sweep is 6 % faster than at the compiler's two wavefronts per SIMD
template <int W, int RK = -1>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(3)))
k_flux_x1(const Inst *__restrict__ I_, Work w) {
  extern __shared__ double lds[];
  WORK_LOOP(w) {
    WORK_INST(w);
    flux_x1_body<W, RK>(g, st, lb, lds);
    __syncthreads();  // psf is rebuilt by the next block of the loop
  }
}

// ---------------------------------------------------------------------------
// Panel-edge L/R exchange (CubedSphere::sendNeighborBlocks, cubed_sphere.cpp:
// 326-576): the interior-side WENO state of each edge face, its IVY / IVZ
// mapped through lat-lon into the neighbour's frame, stored in the
// neighbour's ghost-side slot [n][panel][side][a][i].
__device__ __forceinline__ void edge_body(const DGrid &g, const DState &st, int lb,
                                          const ADDR(1) EdgeDesc *__restrict__ ed, int nedge) {
  int t = lb * BLK + threadIdx.x;
  int ni = g.nx1;
  if (t >= nedge * ni) return;
  int e = t / ni, i = g.is + t % ni;
  EdgeDesc d = *(const EdgeDesc *)(ed + e);
  int p = d.src_panel, k = d.src_k, j = d.src_j;
  double q[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (g.skipv(n)) continue;
    double v;
    if (d.src_dir == 0)  // R state of cell js at face js
      v = weno5(st.w[g.c(n, p, k, i, j - 2)], st.w[g.c(n, p, k, i, j - 1)],
                st.w[g.c(n, p, k, i, j)], st.w[g.c(n, p, k, i, j + 1)],
                st.w[g.c(n, p, k, i, j + 2)]);
    else if (d.src_dir == 1)  // L state of cell je at face je+1
      v = weno5(st.w[g.c(n, p, k, i, j + 1)], st.w[g.c(n, p, k, i, j)],
                st.w[g.c(n, p, k, i, j - 1)], st.w[g.c(n, p, k, i, j - 2)],
                st.w[g.c(n, p, k, i, j - 3)]);
    else if (d.src_dir == 2)
      v = weno5(st.w[g.c(n, p, k - 2, i, j)], st.w[g.c(n, p, k - 1, i, j)],
                st.w[g.c(n, p, k, i, j)], st.w[g.c(n, p, k + 1, i, j)],
                st.w[g.c(n, p, k + 2, i, j)]);
    else
      v = weno5(st.w[g.c(n, p, k + 1, i, j)], st.w[g.c(n, p, k, i, j)],
                st.w[g.c(n, p, k - 1, i, j)], st.w[g.c(n, p, k - 2, i, j)],
                st.w[g.c(n, p, k - 3, i, j)]);
    q[n] = v;
  }
  double U, V, vx, vy;
  vec_rll_from_abp(d.x2s, d.x3s, d.src_block, q[IVY], q[IVZ], &U, &V);
  vec_abp_from_rll(d.x2d, d.x3d, d.tgt_block, U, V, &vx, &vy);
  q[IVY] = vx;
  q[IVZ] = vy;
  size_t base = (((size_t)d.tgt_panel * 4 + d.tgt_dir) * g.N + d.tgt_a) * ni + (i - g.is);
  size_t vs = (size_t)6 * 4 * g.N * ni;
#pragma unroll
  for (int n = 0; n < 5; ++n)
    if (!g.skipv(n)) st.edge[n * vs + base] = q[n];
}

__global__ void __launch_bounds__(BLK) k_edge_states(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    edge_body(g, st, lb, glb(I_[ii_].ed), I_[ii_].ned);
  }
}

// x2 faces (dir = 1) or x3 faces (dir = 2): WENO5 L/R of the face, the panel
// edge's ghost side from the exchanged states, and Hydro::RiemannSolver with
// the cubed-sphere frame maps.  One thread per (panel, k, i, j) face.
// Hydro flux of the x2 (dir 1) or x3 (dir 2) face (p, k, i, j), j = js + fa
// (x2) or k = ks + fa (x3), along-edge index kk: WENO5 L/R states, the
// neighbour panel's state on a panel edge (fa = 0 or N), then riemann().
template <int dir, int EK = -1, int RK = -1>
__device__ __forceinline__ void face_flux64(const DGrid &g, const DState &st, int p, int kk, int fa,
                                            int i, int k, int j, double *fl) {
  constexpr int E = eos_k<EK, RK>();
  int N = g.N, ni = g.nx1;
  double wl[5] = {0, 0, 0, 0, 0}, wr[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (skip_of<E>(g, n)) continue;
    double m3, m2, m1, c0, p1, p2;
    if (dir == 1) {
      m3 = st.w[g.c(n, p, k, i, j - 3)];
      m2 = st.w[g.c(n, p, k, i, j - 2)];
      m1 = st.w[g.c(n, p, k, i, j - 1)];
      c0 = st.w[g.c(n, p, k, i, j)];
      p1 = st.w[g.c(n, p, k, i, j + 1)];
      p2 = st.w[g.c(n, p, k, i, j + 2)];
    } else {
      m3 = st.w[g.c(n, p, k - 3, i, j)];
      m2 = st.w[g.c(n, p, k - 2, i, j)];
      m1 = st.w[g.c(n, p, k - 1, i, j)];
      c0 = st.w[g.c(n, p, k, i, j)];
      p1 = st.w[g.c(n, p, k + 1, i, j)];
      p2 = st.w[g.c(n, p, k + 2, i, j)];
    }
    // on a panel edge the ghost side is the neighbour's state (below)
    if (fa != 0) wl[n] = weno5(p1, c0, m1, m2, m3);
    if (fa != N) wr[n] = weno5(m2, m1, c0, p1, p2);
  }
  // on a panel edge (fa = 0: the L state, fa = N: the R state) the ghost
  // side is the exchanged state; the values are selected, not the array they
  // go to, which keeps wl and wr in registers
  if (fa == 0 || fa == N) {
    size_t vs = (size_t)6 * 4 * N * ni;
    int side = (dir == 1 ? 0 : 2) + (fa == 0 ? 0 : 1);
    size_t b = (((size_t)p * 4 + side) * N + kk) * ni + (i - g.is);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      if (skip_of<E>(g, n)) continue;
      double e = st.edge[n * vs + b];
      wl[n] = fa == 0 ? e : wl[n];
      wr[n] = fa == 0 ? wr[n] : e;
    }
  }
  double cosf, sinf, cth;
  if (dir == 1) {
    cosf = g.T(T_COS2, p, k, j);
    sinf = g.T(T_SIN2, p, k, j);
    cth = g.T(T_CTH2, p, k, j);
  } else {
    cosf = g.T(T_COS3, p, k, j);
    sinf = g.T(T_SIN3, p, k, j);
    cth = g.T(T_CTH3, p, k, j);
  }
  riemann<double>(eos_of<E>(g), rs_of<RK>(g), dir == 1 ? IVY : IVZ, g.gamma, wl, wr, fl, cth, cosf, sinf, 0.);
}

template <int dir, int EK = -1, int RK = -1>
__device__ __forceinline__ void flux_h_body(const DGrid &g, const DState &st, int lb) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  int na = N + 1;  // faces along the normal axis
  size_t total = (size_t)6 * N * ni * na;
  if (t >= total) return;
  // lanes walk j in both directions: x2 faces (j fastest over N+1 faces),
  // x3 faces (j fastest over N cells, then the face index k, then i)
  int kk, fa, i, p;
  if (dir == 1) {
    fa = (int)(t % na);
    size_t r = t / na;
    i = g.is + (int)(r % ni);
    r /= ni;
    kk = (int)(r % N);
    p = (int)(r / N);
  } else {
    // j fastest, then the face index k, then i: the six-row stencil along k
    // of consecutive faces is re-read within a few blocks, from L2
    kk = (int)(t % N);
    size_t r = t / N;
    fa = (int)(r % na);
    r /= na;
    i = g.is + (int)(r % ni);
    p = (int)(r / ni);
  }
  int k, j;
  if (dir == 1) { k = g.ks + kk; j = g.js + fa; }
  else { k = g.ks + fa; j = g.js + kk; }
  double fl[5];
  face_flux64<dir, EK, RK>(g, st, p, kk, fa, i, k, j, fl);
  ADDR(1) double *F = st.flx[dir];
#pragma unroll
  for (int n = 0; n < 5; ++n)
    if (!skip_of<eos_k<EK, RK>()>(g, n)) F[g.f(n, p, k, i, j)] = fl[n];
  for (int n = 0; n < g.nt; ++n) {
    double r[6];
    for (int s = 0; s < 6; ++s)
      r[s] = dir == 1 ? st.r[g.c(n, p, k, i, j - 3 + s)] : st.r[g.c(n, p, k - 3 + s, i, j)];
    double rl = weno5(r[4], r[3], r[2], r[1], r[0]);
    double rr = weno5(r[1], r[2], r[3], r[4], r[5]);
    double m = fl[IDN];
    st.sflx[dir][g.f(n, p, k, i, j)] = m >= 0.0 ? m * rl : m * rr;
  }
}

// five wavefronts per SIMD: 96 VGPRs without spills, 4 % faster than the
// compiler's four on the x3 faces of 6 x 96 x 96 x 64; the shallow-water x2
// faces at six (80 VGPRs without spills, 4 % faster on the 256-instance
// ensemble; its x3 faces spill there).  EK: the equation of state of every
// instance of the launch, or -1 (any); RK: their Riemann solver, or -1.
template <int dir, int EK = -1, int RK = -1>
__global__ void __launch_bounds__(BLK)
__attribute__((amdgpu_waves_per_eu(dir == 1 && EK == EOS_SW ? 6 : 5)))
k_flux_h(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    flux_h_body<dir, EK, RK>(g, st, lb);
  }
}

// Mixed shallow water: the hydro fluxes of the faces on panel edges in FP64
// (face_flux64), so the two panels that share an edge face apply the same
// Riemann flux there, as in the FP64 form.  Threads cover [p][side][kk][i]:
// side 0/1 the x2 faces fa = 0 / N, side 2/3 the x3 faces.
__device__ __forceinline__ void edge_flux_body(const DGrid &g, const DState &st, int lb) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * 4 * N * ni;
  if (t >= total) return;
  int i = g.is + (int)(t % ni);
  size_t r = t / ni;
  int kk = (int)(r % N);
  r /= N;
  int side = (int)(r % 4), p = (int)(r / 4);
  int fa = (side & 1) ? N : 0;
  double fl[5];
  int dir = side < 2 ? 1 : 2;
  int k = dir == 1 ? g.ks + kk : g.ks + fa, j = dir == 1 ? g.js + fa : g.js + kk;
  if (dir == 1) face_flux64<1, EOS_SW>(g, st, p, kk, fa, i, k, j, fl);
  else face_flux64<2, EOS_SW>(g, st, p, kk, fa, i, k, j, fl);
  ADDR(1) double *F = st.flx[dir];
#pragma unroll
  for (int n = 0; n < 5; ++n)
    if (!skip_of<EOS_SW>(g, n)) F[g.f(n, p, k, i, j)] = fl[n];
}

__global__ void __launch_bounds__(BLK) k_edge_flux(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    edge_flux_body(g, st, lb);
  }
}

// ---------------------------------------------------------------------------
// Stage update of one interior cell: SSPRK3 weighted average (u1, u),
// Hydro::AddFluxDivergence, GnomonicEquiangle::AddCoordTermsDivergence
// (gnomonic_equiangle.cpp:834-879), ConstantAcceleration
// (patches/25.constant_acc.patch:19-84) and the Coriolis user source
// (W92.cpp:48-70 / hs94.cpp rotation term).  Shallow water updates u
// directly; the ideal gas writes the tendency du (ImplicitHydroTasks).
struct StageW {
  double delta, g1, g2, beta;
  int first;
  int prim;  // shallow water: also write the primitive state of the cell
  double dfloor, sfloor;
};

// One cell (p, k, i, j) of the stage: the weighted average, AddFluxDivergence,
// the coordinate source terms and the sources.  out[] receives the hydro
// result: the tendency (ideal gas, whose state u is the weighted average and
// is written here) or the new state (shallow water).  Passive scalars are
// updated in place.
// The stage's weighted average of one value of u, formed as update_cell
// forms it (the first stage sets the stage register u1 = delta u and keeps u
// when delta is 1; later stages update u1 when delta != 0); *u1p receives u1
// where update_cell stores it.
__device__ __forceinline__ double stage_avg(const StageW &sw, double u, double u1,
                                            ADDR(1) double *u1p) {
  if (sw.first) {
    double v = 0.;
    if (sw.delta != 0.0) v += sw.delta * u;
    *u1p = v;
    return sw.delta == 1.0 ? u : v;
  }
  if (sw.delta != 0.0) {
    u1 += sw.delta * u;
    *u1p = u1;
  }
  return sw.g1 * u + sw.g2 * u1;
}

// NOAVG (ideal gas within a stage whose column solve follows): the weighted
// average of u is left to the column solve's backward substitution, which
// reads and writes u anyway; here only the tendency is formed.
// NOSCAL: the instance has no passive scalars (the caller checked nt == 0).
// Face fluxes a caller holds in registers: FP64, the cell's x2 faces j and
// j + 1 (x2l, x2r), its x1 faces i and i + 1 (x1l, x1r) and the scalars' x2
// faces (s2l, s2r; up to 4 scalars); mixed, the FP32 x2 faces of the cell and
// of its scalars.
struct RegFlux {
  const double *x2l, *x2r, *x1l, *x1r, *s2l, *s2r;
  const float *f2l, *f2r, *fs2l, *fs2r;
};
constexpr int RF_NT = 4;  // scalars register fluxes carry (the configuration's maximum)

// X2R: the fluxes of rf come from the caller's registers instead of the flux
// arrays (FP64: x2, x1 and the scalars' x2; mixed: the FP32 x2 and scalar x2
// faces).
// cth_out (if not null) receives the cell's cos(theta).
template <bool MIX, bool NOAVG = false, bool NOSCAL = false, bool X2R = false, int EK = -1>
__device__ __forceinline__ void update_cell(const DGrid &g, const DState &st, int p, int k, int i,
                                            int j, StageW sw, double *out,
                                            const RegFlux *rf = nullptr,
                                            double *cth_out = nullptr) {
  double dt = st.dtv[0];
  double wght = sw.beta * dt;
  bool tend = eos_of<EK>(g) == EOS_IDEAL;
  double a1l = 0., a1r = 0.;  // x1 faces: shallow water has no x1 flux
  if (eos_of<EK>(g) != EOS_SW) {
    a1l = g.area1(p, k, j, i);
    a1r = g.area1(p, k, j, i + 1);
  }
  double a2l = g.area2(p, k, j, i), a2r = g.area2(p, k, j + 1, i);
  double a3l = g.area3(p, k, j, i), a3r = g.area3(p, k + 1, j, i);
  // Shallow water has one layer, so per-cell tables are read once per cell (a
  // 3-D column reads them once for all its levels).  Its tangents x, y come
  // from the panel's per-row tables instead, and cos theta = -x y / (C D)
  // from them with the set-up table's expression (exo_setup.cpp), which gives
  // the table's bits for one quotient.
  const bool rows = !X2R && eos_of<EK>(g) == EOS_SW;
  double rx = 0., ry = 0., rcos = 0.;
  if (rows) {
    rx = g.T1(T1_X, p, j);
    ry = g.T1(T1_Y, p, k);
    rcos = qdiv(-rx * ry, g.T1(T1_C, p, j) * g.T1(T1_D, p, k));
  }
  double vol = g.volume(p, k, j, i);
  // every quotient by vol (and by the radius below) shares one correctly
  // rounded reciprocal: div_rn gives the IEEE quotient
  double yvol = qdiv(1., vol);
  // the first stage only sets u1 = delta u (u itself is unchanged when delta
  // is 1), later stages keep u1 unless delta != 0: arrays that would be read
  // only to be overwritten, or rewritten unchanged, are not touched
  bool keep_u = sw.first && sw.delta == 1.0;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (skip_of<EK>(g, n)) continue;
    size_t c = g.c(n, p, k, i, j);
    double u = 0., u1;
    if (NOAVG) {
    } else if (sw.first) {
      u = st.u[c];
      u1 = 0.;
      if (sw.delta != 0.0) u1 += sw.delta * u;
      u = u1;
      st.u1[c] = u1;
    } else {
      u = st.u[c];
      u1 = st.u1[c];
      if (sw.delta != 0.0) {
        u1 += sw.delta * u;
        st.u1[c] = u1;
      }
      u = sw.g1 * u + sw.g2 * u1;
    }
    double dflx;
    if (!MIX) {
      if (eos_of<EK>(g) == EOS_SW) dflx = 0.;
      else
        dflx = a1r * (X2R ? rf->x1r[n] : st.flx[0][g.f(n, p, k, i + 1, j)]) -
               a1l * (X2R ? rf->x1l[n] : st.flx[0][g.f(n, p, k, i, j)]);
      double f2r = X2R ? rf->x2r[n] : st.flx[1][g.f(n, p, k, i, j + 1)];
      double f2l = X2R ? rf->x2l[n] : st.flx[1][g.f(n, p, k, i, j)];
      dflx += a2r * f2r - a2l * f2l;
      dflx += a3r * st.flx[2][g.f(n, p, k + 1, i, j)] - a3l * st.flx[2][g.f(n, p, k, i, j)];
    } else {
      // FP32 perturbation fluxes plus the FP64 reference-pressure parts (the
      // ideal gas; mixed shallow water stores FP64 face fluxes and takes the
      // FP64 update)
      const ADDR(1) float *F0 = st.fflx[0], *F1 = st.fflx[1], *F2 = st.fflx[2];
      dflx = a1r * (double)F0[g.f(n, p, k, i + 1, j)] - a1l * (double)F0[g.f(n, p, k, i, j)];
      dflx += a2r * (double)(X2R ? rf->f2r[n] : F1[g.f(n, p, k, i, j + 1)]) -
              a2l * (double)(X2R ? rf->f2l[n] : F1[g.f(n, p, k, i, j)]);
      dflx += a3r * (double)F2[g.f(n, p, k + 1, i, j)] - a3l * (double)F2[g.f(n, p, k, i, j)];
      double P = st.lref[i];
      if (n == IVX)
        dflx += a1r * st.psf[g.f(0, p, k, i + 1, j)] - a1l * st.psf[g.f(0, p, k, i, j)];
      if (n == IVY)
        dflx += ((a2r * g.T(T_PRY2, p, k, j + 1) - a2l * g.T(T_PRY2, p, k, j)) +
                 (a3r * g.T(T_PRY3, p, k + 1, j) - a3l * g.T(T_PRY3, p, k, j))) * P;
      if (n == IVZ)
        dflx += ((a2r * g.T(T_PRZ2, p, k, j + 1) - a2l * g.T(T_PRZ2, p, k, j)) +
                 (a3r * g.T(T_PRZ3, p, k + 1, j) - a3l * g.T(T_PRZ3, p, k, j))) * P;
    }
    if (tend) {
      if (!keep_u && !NOAVG) st.u[c] = u;
      out[n] = 0.;
      out[n] -= div_rn(wght * dflx, vol, yvol);
    } else {
      out[n] = u;
      out[n] -= div_rn(wght * dflx, vol, yvol);
    }
  }
  // coordinate source terms
  double cosine = rows ? rcos : g.T(T_COSC, p, k, j);
  double sinc = g.T(T_SINC, p, k, j);
  if (cth_out) *cth_out = cosine;
  double sine2 = sinc * sinc;
  // the shallow-water x1 velocity is zero for good (not loaded)
  double v1 = eos_of<EK>(g) == EOS_SW ? 0. : st.w[g.c(IVX, p, k, i, j)], v2 = st.w[g.c(IVY, p, k, i, j)];
  double v3 = st.w[g.c(IVZ, p, k, i, j)];
  double rho = st.w[g.c(IDN, p, k, i, j)];
  double radius = g.x1v[i], yrad = qdiv(1., radius);
  double v_2 = v2 + v3 * cosine, v_3 = v3 + v2 * cosine;
  double pr;
  if (eos_of<EK>(g) == EOS_SW) {
    pr = 0.5 * rho * rho;
  } else {
    pr = st.w[g.c(IPR, p, k, i, j)];
    double src1 = div_rn(2.0 * pr + rho * (v2 * v_2 + v3 * v_3), radius, yrad);
    out[IVX] += wght * src1;
  }
  // x_ov_rD / y_ov_rC of the constructor (gnomonic_equiangle.cpp:209-218)
  double xovrD = div_rn(a2l * g.T(T_SIN2, p, k, j) - a2r * g.T(T_SIN2, p, k, j + 1), vol, yvol);
  double yovrC = div_rn(a3l * g.T(T_SIN3, p, k, j) - a3r * g.T(T_SIN3, p, k + 1, j), vol, yvol);
  double src2 = -xovrD * (pr + rho * v3 * v3 * sine2) - div_rn(rho * v1 * v_2, radius, yrad);
  out[IVY] += wght * src2;
  double src3 = -yovrC * (pr + rho * v2 * v2 * sine2) - div_rn(rho * v1 * v_3, radius, yrad);
  out[IVZ] += wght * src3;
  // sources
  if (eos_of<EK>(g) == EOS_IDEAL && g.grav != 0.0) {
    double src = wght * rho * g.grav;
    out[IVX] += src;
    out[IEN] += src * v1;
  }
  if (g.coriolis != 0) {
    double X = rows ? rx : g.T(T_TANX, p, k, j), Y = rows ? ry : g.T(T_TANY, p, k, j);
    double f = g.T(T_CORF, p, k, j);
    double U, V, acc2, acc3;
    vec_rll_from_abp(X, Y, p + 1, v2, v3, &U, &V);  // CubedSphere::GetUV
    U = -U;
    double llU, llV;
    if (g.coriolis == 1) {
      llU = f * V;
      llV = -f * U;
    } else {
      double m2 = rho * U, m3 = rho * V;
      llU = f * m3;
      llV = -f * m2;
    }
    vec_abp_from_rll(X, Y, p + 1, -llU, llV, &acc2, &acc3);  // GetVyVz
    // (with the row factors: the same C, D and quotient as the cosine above)
    double cth = rows ? rcos : qdiv(-X * Y, sqrt(X * X + 1) * sqrt(Y * Y + 1));
    double a2c = acc2 + acc3 * cth, a3c = acc3 + acc2 * cth;
    if (g.coriolis == 1) {
      out[IVY] += wght * rho * a2c;
      out[IVZ] += wght * rho * a3c;
    } else {
      out[IVY] += wght * a2c;
      out[IVZ] += wght * a3c;
    }
  }
  // passive scalars: weighted average and flux divergence into s (with
  // register fluxes the loop is unrolled over the largest count, so the flux
  // arrays are indexed by constants and stay in registers)
  auto scalar = [&](int n) {
    size_t c = g.c(n, p, k, i, j);
    double s = st.s[c], s1;
    if (sw.first) {
      s1 = 0.;
      if (sw.delta != 0.0) s1 += sw.delta * s;
      s = s1;
      st.s1[c] = s1;
    } else {
      s1 = st.s1[c];
      if (sw.delta != 0.0) {
        s1 += sw.delta * s;
        st.s1[c] = s1;
      }
      s = sw.g1 * s + sw.g2 * s1;
    }
    double dflx;
    if (!MIX) {
      dflx = a1r * st.sflx[0][g.f(n, p, k, i + 1, j)] - a1l * st.sflx[0][g.f(n, p, k, i, j)];
      if (eos_of<EK>(g) == EOS_SW) dflx = 0.;
      double s2r = X2R ? rf->s2r[n] : st.sflx[1][g.f(n, p, k, i, j + 1)];
      double s2l = X2R ? rf->s2l[n] : st.sflx[1][g.f(n, p, k, i, j)];
      dflx += a2r * s2r - a2l * s2l;
      dflx += a3r * st.sflx[2][g.f(n, p, k + 1, i, j)] - a3l * st.sflx[2][g.f(n, p, k, i, j)];
    } else {
      const ADDR(1) float *S0 = st.fsflx[0], *S1 = st.fsflx[1], *S2 = st.fsflx[2];
      dflx = eos_of<EK>(g) == EOS_SW ? 0.
                             : a1r * (double)S0[g.f(n, p, k, i + 1, j)] -
                                   a1l * (double)S0[g.f(n, p, k, i, j)];
      dflx += a2r * (double)(X2R ? rf->fs2r[n] : S1[g.f(n, p, k, i, j + 1)]) -
              a2l * (double)(X2R ? rf->fs2l[n] : S1[g.f(n, p, k, i, j)]);
      dflx += a3r * (double)S2[g.f(n, p, k + 1, i, j)] - a3l * (double)S2[g.f(n, p, k, i, j)];
    }
    s -= div_rn(wght * dflx, vol, yvol);
    st.s[c] = s;
  };
  if constexpr (NOSCAL) {
  } else if constexpr (X2R) {
#pragma unroll
    for (int n = 0; n < RF_NT; ++n)
      if (n < g.nt) scalar(n);
  } else {
    for (int n = 0; n < g.nt; ++n) scalar(n);
  }
}

template <bool MIX, bool NOAVG = false, int EK = -1>
__device__ __forceinline__ void update_body(const DGrid &g, const DState &st, int lb, StageW sw) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  // the validity flag describes the primitive conversion that follows this
  // update (EquationOfState::ConservedToPrimitive resets it on entry,
  // ideal_moist_hydro.cpp:83)
  if (t == 0) *st.bad = 0;
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * N * ni * N;
  if (t >= total) return;
  int j = g.js + (int)(t % N);
  size_t r = t / N;
  int i = g.is + (int)(r % ni);
  r /= ni;
  int k = g.ks + (int)(r % N), p = (int)(r / N);
  double out[5];
  double cth;
  update_cell<MIX, NOAVG, false, false, EK>(g, st, p, k, i, j, sw, out, nullptr, &cth);
  ADDR(1) double *dst = eos_of<EK>(g) == EOS_IDEAL ? st.du : st.u;
  for (int n = 0; n < nv_of<EK>(g); ++n)
    if (!skip_of<EK>(g, n)) dst[g.c(n, p, k, i, j)] = out[n];
  // shallow water within a step: the stage's new state is final here, so its
  // primitive state is written now (k_prim_ghost then covers the ghosts)
  if (sw.prim && eos_of<EK>(g) == EOS_SW)
    prim_from<EK>(g, st, p, k, i, j, out, sw.dfloor, sw.sfloor, cth);
}

// EK: the equation of state of every instance of the launch, or -1 (any)
template <bool MIX, bool NOAVG = false, int EK = -1>
__global__ void __launch_bounds__(BLK) k_update(const Inst *__restrict__ I_, Work w, StageW sw) {
  WORK_LOOP(w) {
    WORK_INST(w);
    update_body<MIX, NOAVG, EK>(g, st, lb, sw);
  }
}

// FP32 flux of the mixed ideal gas at the x2 (dir 1) or x3 (dir 2) face (p,
// k, i, j), j = js + fa (x2) or k = ks + fa (x3), along-edge index kk: WENO5
// of the FP32 perturbation copy fw, the neighbour panel's exchanged FP32
// state on a panel edge, the level's reference density added back and its
// reference pressure P split off (riemann_pert, exo_mixed.hpp).
template <int dir, int RK = -1>
__device__ __forceinline__ void face_flux_mix(const DGrid &g, const DState &st, int p, int kk,
                                              int fa, int i, int k, int j, float *fl) {
#pragma clang fp contract(fast)
  int N = g.N, ni = g.nx1;
  float wl[5] = {0, 0, 0, 0, 0}, wr[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (g.skipv(n)) continue;
    float v[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      int ks = dir == 1 ? k : k - 3 + s, js = dir == 1 ? j - 3 + s : j;
      v[s] = st.fw[g.c(n, p, ks, i, js)];
    }
    wl[n] = weno5f(v[4], v[3], v[2], v[1], v[0]);
    wr[n] = weno5f(v[1], v[2], v[3], v[4], v[5]);
  }
  // panel edges (fa = 0: the L state, fa = N: the R state): the neighbour's
  // state, selected by value so wl and wr stay in registers
  if (fa == 0 || fa == N) {
    size_t vs = (size_t)6 * 4 * N * ni;
    int side = (dir == 1 ? 0 : 2) + (fa == 0 ? 0 : 1);
    size_t b = (((size_t)p * 4 + side) * N + kk) * ni + (i - g.is);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      float e = st.fedge[n * vs + b];
      wl[n] = fa == 0 ? e : wl[n];
      wr[n] = fa == 0 ? wr[n] : e;
    }
  }
  // full density
  float rref = (float)st.lref[2 * g.nc1 + i];
  wl[IDN] += rref;
  wr[IDN] += rref;
  float P = (float)st.lref[g.nc1 + i];
  double cosd, sind, cthd;
  if (dir == 1) {
    cosd = g.T(T_COS2, p, k, j);
    sind = g.T(T_SIN2, p, k, j);
    cthd = g.T(T_CTH2, p, k, j);
  } else {
    cosd = g.T(T_COS3, p, k, j);
    sind = g.T(T_SIN3, p, k, j);
    cthd = g.T(T_CTH3, p, k, j);
  }
  float cosf = (float)cosd, sinf = (float)sind, cth = (float)cthd;
  riemann_pert(EOS_IDEAL, rs_of<RK>(g), dir == 1 ? IVY : IVZ, (float)g.gamma, wl, wr, P, fl, cth, cosf, sinf,
               0.f);
}

// The x2 faces and the stage update (tendency only) of the ideal gas in one
// pass, so the x2 face fluxes never go through memory: FP64, or MIX the mixed
// form (FP32 faces of face_flux_mix, the update of update_cell<true>); SCAL:
// with the passive scalars (their x2 face fluxes too; the update averages s
// in place).
// The pass writes du, s and s1, which no face of the launch reads (the faces
// read w and r; a shallow-water update here would race: it writes the
// primitive state the neighbouring wavefronts' face stencils read).  The
// faces of the instance are numbered row by row (rows (p, k, i) of N + 1
// faces, j fastest); wavefront w takes faces 63 w .. 63 w + 63, lane l face
// 63 w + l.  Lanes 0..62 then update the cell right of their face (its face
// j), with face j + 1 from the next lane; a row's last face updates no cell.
// Every face and every cell is covered once, with the arithmetic of
// face_flux64<1> (or face_flux_mix<1>), the face kernels' scalar fluxes and
// update_cell.
constexpr int F2_FACES = 63;  // cells a wavefront updates (faces per wavefront - 1)

template <bool SCAL, bool MIX, int RK = -1>
__device__ __forceinline__ void face2_update_body(const DGrid &g, const DState &st, int lb,
                                                  StageW sw) {
  if (lb == 0 && threadIdx.x == 0) *st.bad = 0;  // as update_body
  int N = g.N, ni = g.nx1, na = N + 1;
  size_t total = (size_t)6 * N * ni * na;
  int lane = threadIdx.x & 63;
  size_t F = ((size_t)lb * (BLK / 64) + (threadIdx.x >> 6)) * F2_FACES + lane;
  bool act = F < total;
  size_t r = act ? F : total - 1;
  int fa = (int)(r % na);
  r /= na;
  int i = g.is + (int)(r % ni);
  r /= ni;
  int kk = (int)(r % N), p = (int)(r / N);
  int k = g.ks + kk, j = g.js + fa;
  bool cell = act && lane < F2_FACES && fa < N;
  // the update's x1 face fluxes (FP64 form), loaded before the face
  // arithmetic so their latency is spent on it (the x3 ones too would take the
  // kernel from four to three wavefronts per SIMD, which measured slower)
  double f1l[5] = {0., 0., 0., 0., 0.}, f1r[5] = {0., 0., 0., 0., 0.};
  if (!MIX && cell) {
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      f1l[n] = st.flx[0][g.f(n, p, k, i, j)];
      f1r[n] = st.flx[0][g.f(n, p, k, i + 1, j)];
    }
  }
  // the face flux (FP64, or the mixed form's FP32) and the scalars' (WENO5 of
  // r, or of its FP32 copy, across the face, upwinded by the mass flux); every
  // lane takes part in the exchanges
  double fl[5] = {0., 0., 0., 0., 0.}, fr[5] = {0., 0., 0., 0., 0.};
  double sl[RF_NT] = {0., 0., 0., 0.}, sr[RF_NT] = {0., 0., 0., 0.};
  float gl[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, gr[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float tl[RF_NT] = {0.f, 0.f, 0.f, 0.f}, tr[RF_NT] = {0.f, 0.f, 0.f, 0.f};
  if (MIX) {
    if (act) face_flux_mix<1, RK>(g, st, p, kk, fa, i, k, j, gl);
    if (SCAL && act) {
#pragma unroll
      for (int n = 0; n < RF_NT; ++n) {
        if (n >= g.nt) break;
        float q[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) q[s] = st.fw[g.c(g.nv + n, p, k, i, j - 3 + s)];
        float rl = weno5f(q[4], q[3], q[2], q[1], q[0]);
        float rr = weno5f(q[1], q[2], q[3], q[4], q[5]);
        float m = gl[IDN];
        tl[n] = m >= 0.f ? m * rl : m * rr;
      }
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) gr[n] = __shfl_down(gl[n], 1);
    if (SCAL) {
#pragma unroll
      for (int n = 0; n < RF_NT; ++n) tr[n] = __shfl_down(tl[n], 1);
    }
  } else {
    if (act) face_flux64<1, EOS_IDEAL, RK>(g, st, p, kk, fa, i, k, j, fl);
    if (SCAL && act) {
#pragma unroll
      for (int n = 0; n < RF_NT; ++n) {
        if (n >= g.nt) break;
        double q[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) q[s] = st.r[g.c(n, p, k, i, j - 3 + s)];
        double rl = weno5(q[4], q[3], q[2], q[1], q[0]);
        double rr = weno5(q[1], q[2], q[3], q[4], q[5]);
        double m = fl[IDN];
        sl[n] = m >= 0.0 ? m * rl : m * rr;
      }
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) fr[n] = __shfl_down(fl[n], 1);
    if (SCAL) {
#pragma unroll
      for (int n = 0; n < RF_NT; ++n) sr[n] = __shfl_down(sl[n], 1);
    }
  }
  if (cell) {
    double out[5];
    RegFlux rf{fl, fr, f1l, f1r, sl, sr, gl, gr, tl, tr};
    update_cell<MIX, true, !SCAL, true, EOS_IDEAL>(g, st, p, k, i, j, sw, out, &rf);
#pragma unroll
    for (int n = 0; n < 5; ++n) st.du[g.c(n, p, k, i, j)] = out[n];
  }
}

template <bool SCAL, bool MIX, int RK = -1>
__global__ void __launch_bounds__(BLK) k_face2_update(const Inst *__restrict__ I_, Work w,
                                                      StageW sw) {
  WORK_LOOP(w) {
    WORK_INST(w);
    face2_update_body<SCAL, MIX, RK>(g, st, lb, sw);
  }
}

// ---------------------------------------------------------------------------
// ImplicitSolver::FullCorrection on an x1 column (full_correction.cpp:32-165,
// forward_backward.hpp:15-126), one thread per column; the forward sweep keeps
// the previous block row in registers and spills a[i] and delta[i] to the
// coefficient array for the backward substitution.  Ends with u += du.
// FluxJacobian with e = (gm1 + 1) / gm1 * pres / rho given (one division per
// row, carried between levels instead of three 5x5 matrices)
__device__ __forceinline__ M5 jac_e(double gm1, const double *w, double e) {
  double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 / 2 - e) * v1;
  double c2 = e + s2 / 2 - gm1 * v1 * v1;
  M5 J = {{{0, 1., 0., 0., 0.},
           {gm1 * s2 / 2 - v1 * v1, (2. - gm1) * v1, -gm1 * v2, -gm1 * v3, gm1},
           {-v1 * v2, v2, v1, 0., 0.},
           {-v1 * v3, v3, 0., v1, 0.},
           {c1, c2, -gm1 * v2 * v1, -gm1 * v3 * v1, (gm1 + 1) * v1}}};
  return J;
}

constexpr int IMPL_MAXNC1 = 512;  // radial cells incl. ghosts the column solve supports

// geo: LDS for the instance's x1f | x1v | dx1f, so the per-level geometry is
// read without waiting on the vector-memory counter (which also counts the
// coefficient stores and the prefetches in flight)
// AVG (within a stage): the update kernel formed only the tendency, and the
// stage's weighted average of u (with the stage register u1) is formed here,
// as update_cell forms it, before u += delta: u is read and written once per
// stage instead of twice.
typedef double dpair __attribute__((ext_vector_type(2)));

// Column access of the 5x5 solve on the handle's grid: primitives projected
// with the column's cos/sin (solve_implicit_3d.cpp:117-123), the tendency du,
// the x1 face areas and volume of each level, and the end of the backward
// substitution: u (or, with AVG, the stage's weighted average of u and u1)
// plus the correction.
template <bool AVG>
struct ColIOGrid {
  const ADDR(1) double *__restrict__ W;
  const ADDR(1) double *__restrict__ DU;
  ADDR(1) double *__restrict__ U;
  ADDR(1) double *__restrict__ U1;
  size_t vs, jp, c0;
  double cth, sth, a23, dx2a, dx3a;
  const double *x1f, *x1v, *dx1f;
  StageW sw;
  bool first;
  __device__ __forceinline__ void wraw(int i, double *w) const {
#pragma unroll
    for (int n = 0; n < 5; ++n) w[n] = W[n * vs + c0 + (size_t)i * jp];
  }
  __device__ __forceinline__ void project(double *w) const {
    w[IVY] += w[IVZ] * cth;
    w[IVZ] *= sth;
  }
  __device__ __forceinline__ double du(int i, int n) const {
    return DU[n * vs + c0 + (size_t)i * jp];
  }
  // Grid::area1 = ((DX2A DX3A) x1f) x1f SINC and Grid::volume =
  // ((((dx1f DX2A) DX3A) x1v) x1v) SINC with the same roundings
  __device__ __forceinline__ void geom(int i, double &aleft, double &aright, double &v) const {
    aleft = ((a23 * x1f[i]) * x1f[i]) * sth;
    aright = ((a23 * x1f[i + 1]) * x1f[i + 1]) * sth;
    v = ((((dx1f[i] * dx2a) * dx3a) * x1v[i]) * x1v[i]) * sth;
  }
  __device__ __forceinline__ void uload(int l, V5 &u, V5 &u1) const {
#pragma unroll
    for (int n = 0; n < 5; ++n) u.a[n] = U[n * vs + c0 + (size_t)l * jp];
    if (AVG && !first) {
#pragma unroll
      for (int n = 0; n < 5; ++n) u1.a[n] = U1[n * vs + c0 + (size_t)l * jp];
    }
  }
  // only u += delta is stored: nothing reads the correction itself afterwards
  __device__ __forceinline__ void finish(int l, const V5 &uo, const V5 &u1o, const V5 &d) const {
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      size_t c = n * vs + c0 + (size_t)l * jp;
      double ub = AVG ? stage_avg(sw, uo.a[n], first ? 0. : u1o.a[n], U1 + c) : uo.a[n];
      U[c] = ub + d.a[n];
    }
  }
};

// Column access of the single-column shim: projected primitives w[n][i] and
// the tendency du[n][i] of nc1 levels (x1 ghosts included), the x1 face areas
// of the interior faces and the interior volumes; the correction itself is
// written back into du.
struct ColIOShim {
  const double *w;
  double *dun;
  const double *area, *vol;
  int nc1, is;
  __device__ __forceinline__ void wraw(int i, double *wv) const {
#pragma unroll
    for (int n = 0; n < 5; ++n) wv[n] = w[n * nc1 + i];
  }
  __device__ __forceinline__ void project(double *) const {}
  __device__ __forceinline__ double du(int i, int n) const { return dun[n * nc1 + i]; }
  __device__ __forceinline__ void geom(int i, double &aleft, double &aright, double &v) const {
    aleft = area[i - is];
    aright = area[i - is + 1];
    v = vol[i - is];
  }
  __device__ __forceinline__ void uload(int, V5 &, V5 &) const {}
  __device__ __forceinline__ void finish(int l, const V5 &, const V5 &, const V5 &d) const {
#pragma unroll
    for (int n = 0; n < 5; ++n) dun[n * nc1 + l] = d.a[n];
  }
};

// The 5x5 column solve (full_correction.cpp:32-165, forward_backward.hpp:
// 15-126) over levels is..ie of one column, with time step dt.  CFP is the
// lane's base of the coefficient scratch, [level][pair][64 lanes]: the 30
// doubles of a level (a^-1 c row-major, then delta) as 15 pairs, each pair 64
// lanes x 16 B contiguous, so a level is 15 coalesced 16-byte stores, and 15
// loads back.  Both are non-temporal: each coefficient is written once and read
// once a whole sweep later, and streaming it past the caches leaves them to
// the state (the solve 0.61 -> 0.56 ms on 6 x 96 x 96 x 64).
template <bool AVG, typename IO>
__device__ __forceinline__ void col5_solve(const IO &io, int is, int ie, double dt, double gamma,
                                           double grav, ADDR(1) dpair *__restrict__ CFP) {
  auto cfp = [&](int i, int q2) -> ADDR(1) dpair * {
    return CFP + ((size_t)(i - is) * 15 + q2) * 64;
  };
  double gm1c = (gamma - 1.) * 1. / 1.;
  double gm1 = 0.5 * (gm1c + gm1c);
  double k1 = qdiv(gm1c + 1, gm1c);
  auto erow = [&](const double *w) { return qdiv(k1 * w[IPR], w[IDN]); };
  // |A| at face is-1/2: the reference's row is-1 (full_correction.cpp:272-325)
  // leaves Am = R|L|R^-1 of the Roe state of (w(is-1), w(is))
  double wm[5], w0[5], wp[5], prim[5];
  io.wraw(is - 1, wm);
  io.project(wm);
  io.wraw(is, w0);
  io.project(w0);
  roe_average(prim, gm1, wm, w0);
  double cs = sqrt(qdiv((1. + (gamma - 1) * 1. / 1.) * prim[IPR], prim[IDN]));
  M5 Am = abs_a(prim, cs, gm1);
  double em = erow(wm), e0 = erow(w0);
  // divisions by dt and by 2 vol share one correctly rounded reciprocal each
  double ydt = qdiv(1., dt);
  M5 Aprev;
  V5 dprev;
  // The next level's state and tendency are loaded one level ahead and put
  // to use (projected, divided by dt) before the level's coefficient stores
  // are issued, so a level never waits for the previous level's stores.
  double wnext[5], rnext[5], enext;
  io.wraw(is + 1, wnext);
  io.project(wnext);
  enext = erow(wnext);
#pragma unroll
  for (int n = 0; n < 5; ++n) rnext[n] = div_rn(io.du(is, n), dt, ydt);
  // One level of the forward sweep.  The first and the last level are
  // separate instances (FIRST: no previous row and the reflecting bottom fold;
  // LAST: no prefetch and the reflecting top fold), so the levels between run
  // without selects, and c is formed after T there, which keeps fewer 5x5
  // matrices live at once.
  auto level = [&](auto first_t, auto last_t, int i) {
    constexpr bool FIRST = decltype(first_t)::value, LAST = decltype(last_t)::value;
    double rhsv[5], wnx[5], dun[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      wp[n] = wnext[n];
      rhsv[n] = rnext[n];
    }
    double ep = enext;
    if (!LAST) {  // prefetch the next level
      io.wraw(i + 2, wnx);
#pragma unroll
      for (int n = 0; n < 5; ++n) dun[n] = io.du(i + 1, n);
    }
    double aleft, aright, v;
    io.geom(i, aleft, aright, v);
    double v2 = 2. * v, yv2 = qdiv(1., v2);
    // b, then b Aprev and b dprev, so b dies early (the first level keeps it
    // for the bottom fold)
    M5 b;
    {
      M5 Jm = jac_e(gm1c, wm, em);
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int q = 0; q < 5; ++q)
          b.a[r][q] = div_rn(-(Am.a[r][q] + Jm.a[r][q]) * aleft, v2, yv2);
    }
    M5 ba;
    V5 bd;
    if (!FIRST) {
      ba = m5_mul(b, Aprev);
      bd = m5_mulv(b, dprev);
    }
    roe_average(prim, gm1, w0, wp);
    cs = sqrt(qdiv((1. + (gamma - 1) * 1. / 1.) * prim[IPR], prim[IDN]));
    M5 Ap = abs_a(prim, cs, gm1);
    auto cmat = [&]() {
      M5 c;
      M5 Jp = jac_e(gm1c, wp, ep);
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int q = 0; q < 5; ++q)
          c.a[r][q] = div_rn(-(Ap.a[r][q] - Jp.a[r][q]) * aright, v2, yv2);
      return c;
    };
    M5 c;
    if (LAST) c = cmat();  // the top fold needs c in T
    M5 T;
    {
      M5 J0 = jac_e(gm1c, w0, e0);
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          double dtq = r == q ? ydt : 0.;
          double phi = (r == IVX && q == IDN) || (r == IEN && q == IVX) ? grav : 0.;
          double a = div_rn(Am.a[r][q] * aleft + Ap.a[r][q] * aright +
                                (aright - aleft) * J0.a[r][q],
                            v2, yv2) +
                     dtq - phi;
          if (FIRST) a += b.a[r][q] * (q == IVX ? -1. : 1.);  // reflecting bottom
          if (LAST) a += c.a[r][q] * (q == IVX ? -1. : 1.);   // reflecting top
          T.a[r][q] = FIRST ? a : a - ba.a[r][q];
        }
    }
    if (!LAST) c = cmat();
    V5 tv;
#pragma unroll
    for (int n = 0; n < 5; ++n) tv.a[n] = FIRST ? rhsv[n] : rhsv[n] - bd.a[n];
    M5 ainv = m5_inv(T);
    V5 delta = m5_mulv(ainv, tv);
    Aprev = m5_mul(ainv, c);
    dprev = delta;
    if (!LAST) {  // use the prefetched next level before this level's stores
#pragma unroll
      for (int n = 0; n < 5; ++n) wnext[n] = wnx[n];
      io.project(wnext);
      enext = erow(wnext);
#pragma unroll
      for (int n = 0; n < 5; ++n) rnext[n] = div_rn(dun[n], dt, ydt);
    }
    {
      double cv[30];
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int q = 0; q < 5; ++q) cv[r * 5 + q] = Aprev.a[r][q];
#pragma unroll
      for (int n = 0; n < 5; ++n) cv[25 + n] = delta.a[n];
#pragma unroll
      for (int q2 = 0; q2 < 15; ++q2)
        __builtin_nontemporal_store(dpair{cv[2 * q2], cv[2 * q2 + 1]}, cfp(i, q2));
    }
    Am = Ap;
    em = e0;
    e0 = ep;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      wm[n] = w0[n];
      w0[n] = wp[n];
    }
  };
  using T1 = std::integral_constant<bool, true>;
  using F0 = std::integral_constant<bool, false>;
  if (is == ie) {
    level(T1{}, T1{}, is);
  } else {
    level(T1{}, F0{}, is);
    for (int i = is + 1; i < ie; ++i) level(F0{}, F0{}, i);
    level(F0{}, T1{}, ie);
  }
  // backward substitution; delta(ie) is final
  V5 dnext = dprev;
  {
    V5 uo, u1o;
    io.uload(ie, uo, u1o);
    io.finish(ie, uo, u1o, dnext);
  }
  // levels are loaded ahead within the 63 loads the vector-memory counter
  // tracks: three levels of 20 loads, or two of 25 with the stage register
  constexpr int PD = AVG ? 2 : 3;
  M5 Aq[PD];
  V5 dq[PD], uq[PD], u1q[PD];
  auto bload = [&](int l, M5 &A, V5 &d, V5 &u, V5 &u1) {
    double cv[30];
#pragma unroll
    for (int q2 = 0; q2 < 15; ++q2) {
      dpair t = __builtin_nontemporal_load(cfp(l, q2));
      cv[2 * q2] = t.x;
      cv[2 * q2 + 1] = t.y;
    }
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
      for (int q = 0; q < 5; ++q) A.a[r][q] = cv[r * 5 + q];
#pragma unroll
    for (int n = 0; n < 5; ++n) d.a[n] = cv[25 + n];
    io.uload(l, u, u1);
  };
#pragma unroll
  for (int q = 0; q < PD; ++q)
    if (ie - 1 - q >= is) bload(ie - 1 - q, Aq[q], dq[q], uq[q], u1q[q]);
  for (int i0 = ie - 1; i0 >= is; i0 -= PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      int i = i0 - q;
      if (i < is) break;
      M5 A = Aq[q];
      V5 d = dq[q], uo = uq[q], u1o = u1q[q];
      if (i - PD >= is) bload(i - PD, Aq[q], dq[q], uq[q], u1q[q]);
      V5 ad = m5_mulv(A, dnext);
#pragma unroll
      for (int n = 0; n < 5; ++n) d.a[n] = d.a[n] - ad.a[n];
      io.finish(i, uo, u1o, d);
      dnext = d;
    }
  }
}

template <bool AVG>
__device__ __forceinline__ void implicit_body(const DGrid &g, const DState &st, int lb,
                                              const StageW &sw, double *geo) {
  for (int q = threadIdx.x; q <= g.nc1; q += blockDim.x) {
    geo[q] = g.x1f[q];
    if (q < g.nc1) {
      geo[IMPL_MAXNC1 + 1 + q] = g.x1v[q];
      geo[2 * (IMPL_MAXNC1 + 1) + q] = g.dx1f[q];
    }
  }
  __syncthreads();
  int N = g.N;
  int nbx = (N * N + 63) / 64;
  int col = (lb % nbx) * 64 + threadIdx.x;
  int p = lb / nbx;
  if (col >= N * N) return;
  int k = g.ks + col / N, j = g.js + col % N;
  ColIOGrid<AVG> io;
  io.W = st.w;
  io.DU = st.du;
  io.U = st.u;
  io.U1 = st.u1;
  io.vs = g.vstride;
  io.jp = g.jp;
  io.c0 = g.c(0, p, k, 0, j);  // column base (variable 0, i = 0)
  io.cth = g.T(T_COSC, p, k, j);
  io.sth = g.T(T_SINC, p, k, j);
  io.dx2a = g.T(T_DX2A, p, k, j);
  io.dx3a = g.T(T_DX3A, p, k, j);
  io.a23 = io.dx2a * io.dx3a;
  io.x1f = geo;
  io.x1v = geo + IMPL_MAXNC1 + 1;
  io.dx1f = geo + 2 * (IMPL_MAXNC1 + 1);
  io.sw = sw;
  io.first = sw.first;
  ADDR(1) dpair *CFP = (ADDR(1) dpair *)st.coef + (size_t)lb * g.nx1 * 15 * 64 + threadIdx.x;
  col5_solve<AVG>(io, g.is, g.ie, sw.beta * st.dtv[0], g.gamma, g.grav, CFP);
}

template <bool AVG>
__global__ void __launch_bounds__(64) k_implicit(const Inst *__restrict__ I_, Work w, StageW sw) {
  __shared__ double geo[3 * (IMPL_MAXNC1 + 1)];
  WORK_LOOP(w) {
    WORK_INST(w);
    implicit_body<AVG>(g, st, lb, sw, geo);
    __syncthreads();  // geo is refilled by the next block of the loop
  }
}

// ncol columns through the same solve (the column shims), one lane per
// column: column c's arrays follow column c - 1's (w and du nc1 levels of 5
// variables, nx1 + 1 face areas, nx1 volumes); coef holds the coefficient
// scratch, [64-column block][level][pair][64 lanes], written by each column's
// forward sweep before its backward substitution reads it.
__global__ void __launch_bounds__(64) k_implicit_columns(ColIOShim io, int ncol, int nx1,
                                                         double dt, double gamma, double grav,
                                                         double *coef) {
  int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= ncol) return;
  ColIOShim cio = io;
  cio.w += (size_t)c * 5 * io.nc1;
  cio.dun += (size_t)c * 5 * io.nc1;
  cio.area += (size_t)c * (nx1 + 1);
  cio.vol += (size_t)c * nx1;
  ADDR(1) dpair *CFP =
      (ADDR(1) dpair *)glb(coef) + (size_t)blockIdx.x * nx1 * 15 * 64 + threadIdx.x;
  col5_solve<false>(cio, NG, NG + nx1 - 1, dt, gamma, grav, CFP);
}

// ImplicitSolver::PartialCorrection on an x1 column (partial_correction.cpp:
// 26-177): the block-tridiagonal system of the (density, x1 momentum, energy)
// rows and columns of the same |A| and flux Jacobians, gravity in Phi, the
// reflecting ends diag(1, -1, 1).  Only those three tendencies are replaced;
// the transverse momenta take their explicit tendency.  Ends with u += du.
constexpr int SUB3[3] = {IDN, IVX, IEN};

template <bool AVG>
__device__ __forceinline__ void implicit3_body(const DGrid &g, const DState &st, int lb,
                                               const StageW &sw, double *geo) {
  double beta = sw.beta;
  for (int q = threadIdx.x; q <= g.nc1; q += blockDim.x) {
    geo[q] = g.x1f[q];
    if (q < g.nc1) {
      geo[IMPL_MAXNC1 + 1 + q] = g.x1v[q];
      geo[2 * (IMPL_MAXNC1 + 1) + q] = g.dx1f[q];
    }
  }
  __syncthreads();
  int N = g.N;
  int nbx = (N * N + 63) / 64;
  int col = (lb % nbx) * 64 + threadIdx.x;
  int p = lb / nbx;
  if (col >= N * N) return;
  int k = g.ks + col / N, j = g.js + col % N;
  int is = g.is, ie = g.ie;
  double dt = beta * st.dtv[0];
  double gamma = g.gamma, grav = g.grav;
  double cth = g.T(T_COSC, p, k, j), sth = g.T(T_SINC, p, k, j);
  double dx2a = g.T(T_DX2A, p, k, j), dx3a = g.T(T_DX3A, p, k, j);
  double a23 = dx2a * dx3a;
  const double *x1f = geo, *x1v = geo + IMPL_MAXNC1 + 1, *dx1f = geo + 2 * (IMPL_MAXNC1 + 1);
  const ADDR(1) double *__restrict__ W = st.w;
  const ADDR(1) double *__restrict__ DU = st.du;
  ADDR(1) double *__restrict__ U = st.u;
  // coefficient scratch, [block][level][pair][lane]: the 12 doubles of a level
  // (a^-1 c row-major, then delta) as 6 pairs of 64 lanes x 16 B
  typedef double dpair __attribute__((ext_vector_type(2)));
  ADDR(1) dpair *__restrict__ CFP =
      (ADDR(1) dpair *)st.coef + (size_t)lb * g.nx1 * 6 * 64 + threadIdx.x;
  auto cfp = [&](int i, int q2) -> ADDR(1) dpair * {
    return CFP + ((size_t)(i - is) * 6 + q2) * 64;
  };
  size_t vs = g.vstride, jp = g.jp;
  size_t c0 = g.c(0, p, k, 0, j);
  auto wload = [&](int i, double *w) {
#pragma unroll
    for (int n = 0; n < 5; ++n) w[n] = W[n * vs + c0 + (size_t)i * jp];
  };
  auto wproj = [&](double *w) {
    w[IVY] += w[IVZ] * cth;
    w[IVZ] *= sth;
  };
  double gm1c = (gamma - 1.) * 1. / 1.;
  double gm1 = 0.5 * (gm1c + gm1c);
  double k1 = qdiv(gm1c + 1, gm1c);
  auto erow = [&](const double *w) { return qdiv(k1 * w[IPR], w[IDN]); };
  double wm[5], w0[5], wp[5], prim[5];
  wload(is - 1, wm);
  wproj(wm);
  wload(is, w0);
  wproj(w0);
  roe_average(prim, gm1, wm, w0);
  double cs = sqrt(qdiv((1. + (gamma - 1) * 1. / 1.) * prim[IPR], prim[IDN]));
  M5 Am = abs_a(prim, cs, gm1);
  double em = erow(wm), e0 = erow(w0);
  double ydt = qdiv(1., dt);
  M3 Aprev;
  V3 dprev;
  // the next level's state and tendency are loaded a level ahead and put to
  // use before the level's coefficient stores
  double wnext[5], enext;
  V3 rnext;
  wload(is + 1, wnext);
  wproj(wnext);
  enext = erow(wnext);
#pragma unroll
  for (int n = 0; n < 3; ++n) rnext.a[n] = div_rn(DU[SUB3[n] * vs + c0 + (size_t)is * jp], dt, ydt);
  for (int i = is; i <= ie; ++i) {
    double wnx[5], dun[3];
#pragma unroll
    for (int n = 0; n < 5; ++n) wp[n] = wnext[n];
    double ep = enext;
    V3 rhs = rnext;
    if (i < ie) {
      wload(i + 2, wnx);
#pragma unroll
      for (int n = 0; n < 3; ++n) dun[n] = DU[SUB3[n] * vs + c0 + (size_t)(i + 1) * jp];
    }
    double aleft = ((a23 * x1f[i]) * x1f[i]) * sth;
    double aright = ((a23 * x1f[i + 1]) * x1f[i + 1]) * sth;
    double v = ((((dx1f[i] * dx2a) * dx3a) * x1v[i]) * x1v[i]) * sth;
    double v2 = 2. * v, yv2 = qdiv(1., v2);
    roe_average(prim, gm1, w0, wp);
    cs = sqrt(qdiv((1. + (gamma - 1) * 1. / 1.) * prim[IPR], prim[IDN]));
    M5 Ap = abs_a(prim, cs, gm1);
    M5 Jm = jac_e(gm1c, wm, em), J0 = jac_e(gm1c, w0, e0), Jp = jac_e(gm1c, wp, ep);
    M3 b, c, T;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        int R = SUB3[r], Q = SUB3[q];
        b.a[r][q] = div_rn(-(Am.a[R][Q] + Jm.a[R][Q]) * aleft, v2, yv2);
        c.a[r][q] = div_rn(-(Ap.a[R][Q] - Jp.a[R][Q]) * aright, v2, yv2);
      }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        int R = SUB3[r], Q = SUB3[q];
        double dtq = r == q ? ydt : 0.;
        double phi = (r == 1 && q == 0) || (r == 2 && q == 1) ? grav : 0.;
        double a = div_rn(Am.a[R][Q] * aleft + Ap.a[R][Q] * aright +
                              (aright - aleft) * J0.a[R][Q],
                          v2, yv2) +
                   dtq - phi;
        if (i == is) a += b.a[r][q] * (q == 1 ? -1. : 1.);  // reflecting bottom
        if (i == ie) a += c.a[r][q] * (q == 1 ? -1. : 1.);  // reflecting top
        T.a[r][q] = a;
      }
    V3 tv = rhs;
    if (i > is) {
      M3 ba = m3_mul(b, Aprev);
      V3 bd = m3_mulv(b, dprev);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int q = 0; q < 3; ++q) T.a[r][q] = T.a[r][q] - ba.a[r][q];
        tv.a[r] = rhs.a[r] - bd.a[r];
      }
    }
    M3 ainv = m3_inv(T);
    V3 delta = m3_mulv(ainv, tv);
    Aprev = m3_mul(ainv, c);
    dprev = delta;
    if (i < ie) {
#pragma unroll
      for (int n = 0; n < 5; ++n) wnext[n] = wnx[n];
      wproj(wnext);
      enext = erow(wnext);
#pragma unroll
      for (int n = 0; n < 3; ++n) rnext.a[n] = div_rn(dun[n], dt, ydt);
    }
    double cv[12];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) cv[r * 3 + q] = Aprev.a[r][q];
#pragma unroll
    for (int n = 0; n < 3; ++n) cv[9 + n] = delta.a[n];
#pragma unroll
    for (int q2 = 0; q2 < 6; ++q2)
      __builtin_nontemporal_store(dpair{cv[2 * q2], cv[2 * q2 + 1]}, cfp(i, q2));
    Am = Ap;
    em = e0;
    e0 = ep;
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      wm[n] = w0[n];
      w0[n] = wp[n];
    }
  }
  // backward substitution, then u += du: the corrected (density, x1 momentum,
  // energy) and the explicit transverse momenta
  // each level reads 6 coefficient pieces, 5 states, 2 tendencies and (AVG,
  // after the first stage) 5 stage registers; levels are loaded four ahead
  // (52 loads in flight), or three (54) with the stage registers, within the
  // 63 the vector-memory counter tracks.  AVG (within a stage): delta is
  // added to the stage's weighted average of u, formed here.
  ADDR(1) double *U1 = st.u1;
  auto ustore = [&](int i, const V3 &d, const double *uo) {
    size_t ci = c0 + (size_t)i * jp;
    double ub[5];
#pragma unroll
    for (int n = 0; n < 5; ++n)
      ub[n] = AVG ? stage_avg(sw, uo[n], sw.first ? 0. : uo[5 + n], U1 + n * vs + ci) : uo[n];
#pragma unroll
    for (int n = 0; n < 3; ++n) U[SUB3[n] * vs + ci] = ub[SUB3[n]] + d.a[n];
    U[IVY * vs + ci] = ub[IVY] + uo[10];
    U[IVZ * vs + ci] = ub[IVZ] + uo[11];
  };
  // uo: u (5), the stage register u1 (5), du(IVY), du(IVZ)
  auto uload = [&](int i, double *uo) {
    size_t ci = c0 + (size_t)i * jp;
#pragma unroll
    for (int n = 0; n < 5; ++n) uo[n] = U[n * vs + ci];
    if (AVG && !sw.first) {
#pragma unroll
      for (int n = 0; n < 5; ++n) uo[5 + n] = U1[n * vs + ci];
    }
    uo[10] = DU[IVY * vs + ci];
    uo[11] = DU[IVZ * vs + ci];
  };
  constexpr int PD = AVG ? 3 : 4;
  double cq[PD][12], uq[PD][12];
  auto bload = [&](int l, double *cv, double *uo) {
#pragma unroll
    for (int q2 = 0; q2 < 6; ++q2) {
      dpair t = __builtin_nontemporal_load(cfp(l, q2));
      cv[2 * q2] = t.x;
      cv[2 * q2 + 1] = t.y;
    }
    uload(l, uo);
  };
  V3 dnext = dprev;
  {
    double uo[12];
    uload(ie, uo);
    ustore(ie, dnext, uo);
  }
#pragma unroll
  for (int q = 0; q < PD; ++q)
    if (ie - 1 - q >= is) bload(ie - 1 - q, cq[q], uq[q]);
  for (int i0 = ie - 1; i0 >= is; i0 -= PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      int i = i0 - q;
      if (i < is) break;
      double cv[12], uo[12];
#pragma unroll
      for (int t = 0; t < 12; ++t) cv[t] = cq[q][t];
#pragma unroll
      for (int t = 0; t < 12; ++t) uo[t] = uq[q][t];
      if (i - PD >= is) bload(i - PD, cq[q], uq[q]);
      M3 A;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) A.a[r][c] = cv[r * 3 + c];
      V3 ad = m3_mulv(A, dnext);
      V3 d;
#pragma unroll
      for (int n = 0; n < 3; ++n) d.a[n] = cv[9 + n] - ad.a[n];
      ustore(i, d, uo);
      dnext = d;
    }
  }
}

template <bool AVG>
__global__ void __launch_bounds__(64) k_implicit3(const Inst *__restrict__ I_, Work w,
                                                  StageW sw) {
  __shared__ double geo[3 * (IMPL_MAXNC1 + 1)];
  WORK_LOOP(w) {
    WORK_INST(w);
    implicit3_body<AVG>(g, st, lb, sw, geo);
    __syncthreads();
  }
}


// u += du without the column solve (ImplicitHydroTasks with implicit_flag 0)
// AVG (within a stage): u = the stage's weighted average + du
template <bool AVG>
__device__ __forceinline__ void add_du_body(const DGrid &g, const DState &st, int lb,
                                            const StageW &sw) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * N * ni * N;
  if (t >= total) return;
  int j = g.js + (int)(t % N);
  size_t r = t / N;
  int i = g.is + (int)(r % ni);
  r /= ni;
  int k = g.ks + (int)(r % N), p = (int)(r / N);
  for (int n = 0; n < g.nv; ++n) {
    size_t c = g.c(n, p, k, i, j);
    double u = st.u[c];
    if (AVG) u = stage_avg(sw, u, sw.first ? 0. : st.u1[c], st.u1 + c);
    st.u[c] = u + st.du[c];
  }
}

template <bool AVG>
__global__ void __launch_bounds__(BLK) k_add_du(const Inst *__restrict__ I_, Work w, StageW sw) {
  WORK_LOOP(w) {
    WORK_INST(w);
    add_du_body<AVG>(g, st, lb, sw);
  }
}

// ---------------------------------------------------------------------------
// Hydro::NewBlockTimeStep (patches/21.new_blockdt.patch:16-37): per-cell
// width / signal speed, block minimum, device-wide minimum on the IEEE bits
// (positive doubles order like unsigned integers).
// one block per instance (work list of one block each), thread 0 acts
__global__ void k_dt_reset(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    (void)g;
    (void)lb;
    if (threadIdx.x == 0) *st.dtbits = 0x7FF0000000000000ull;
  }
}

constexpr int DTC = 8;  // cells per thread of the CFL minimum (fewer block atomics)

__device__ __forceinline__ void dt_body(const DGrid &g, const DState &st, int lb, double *wmin) {
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * N * ni * N;
  double m = HUGE_VAL;
  for (int cc = 0; cc < DTC; ++cc) {
    size_t t = ((size_t)lb * DTC + cc) * BLK + threadIdx.x;
    if (t >= total) break;
    // 32-bit quotients (an instance has fewer than 2^32 cells, which
    // launcher_setup checks): 0.044 -> 0.042 ms on 6 x 96 x 96 x 64
    unsigned tu = (unsigned)t, Nu = (unsigned)N, niu = (unsigned)ni;
    int j = g.js + (int)(tu % Nu);
    unsigned r = tu / Nu;
    int i = g.is + (int)(r % niu);
    r /= niu;
    int k = g.ks + (int)(r % Nu), p = (int)(r / Nu);
    double wi[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (!g.skipv(n)) wi[n] = st.w[g.c(n, p, k, i, j)];
    double cs = g.eos == EOS_SW ? sqrt(wi[IDN])
                                : sqrt(qdiv((1. + (g.gamma - 1) * 1. / 1.) * wi[IPR], wi[IDN]));
    double dt1 = g.dx1f[i];
    double dt2 = g.x1v[i] * g.T(T_DX2A, p, k, j), dt3 = g.x1v[i] * g.T(T_DX3A, p, k, j);
    double sp1 = fabs(wi[IVX]) + cs, sp2 = fabs(wi[IVY]) + cs, sp3 = fabs(wi[IVZ]) + cs;
    if ((g.implicit & 1) && N > 1) dt1 = qdiv(dt1, g.cfl * fabs(wi[IVX]));
    else dt1 = qdiv(dt1, sp1);
    dt2 = qdiv(dt2, sp2);
    dt3 = qdiv(dt3, sp3);
    if (g.cfl_x1 && dt1 < m) m = dt1;
    if (dt2 < m) m = dt2;
    if (dt3 < m) m = dt3;
  }
  // wavefront then block minimum
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off);
    m = o < m ? o : m;
  }
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) wmin[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wmin[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) b = wmin[q] < b ? wmin[q] : b;
    atomicMin((unsigned long long *)st.dtbits, (unsigned long long)__double_as_longlong(b));
  }
}

__global__ void __launch_bounds__(BLK) k_dt(const Inst *__restrict__ I_, Work w) {
  __shared__ double wmin[BLK / 64];
  WORK_LOOP(w) {
    WORK_INST(w);
    dt_body(g, st, lb, wmin);
    __syncthreads();  // wmin is reused by the next block of the loop
  }
}

// dt = min(cfl * min_cells, 2 dt_prev) (Mesh::NewTimeStep growth limit)
__global__ void k_dt_finish(const Inst *__restrict__ I_, Work w, double dt_fixed) {
  WORK_LOOP(w) {
    WORK_INST(w);
    (void)lb;
    if (threadIdx.x != 0) continue;
    if (dt_fixed > 0.0) {
      st.dtv[0] = dt_fixed;
      continue;
    }
    double d = __longlong_as_double((long long)*st.dtbits) * g.cfl;
    double dt = 2.0 * st.dtv[1];
    if (d < dt) dt = d;
    st.dtv[0] = dt;
    st.dtv[1] = dt;
  }
}

// ---------------------------------------------------------------------------
// Conserved totals sum_cells u_n * volume in FP64 with compensated (TwoSum /
// Neumaier) addition: per-block (sum, error) pairs in a fixed tree order, then
// one block per instance adds its partials in a fixed order, so the totals are
// deterministic and accurate to about one rounding of the result.
struct Pair {
  double s, c;
};

__device__ __forceinline__ Pair padd(Pair a, Pair b) {
  double s = a.s + b.s;
  double bb = s - a.s;
  double err = (a.s - (s - bb)) + (b.s - bb);  // exact error of a.s + b.s
  return Pair{s, a.c + b.c + err};
}

__device__ __forceinline__ Pair wave_sum(Pair p) {
  for (int off = 32; off > 0; off >>= 1) {
    Pair o{__shfl_xor(p.s, off), __shfl_xor(p.c, off)};
    // the lower lane of each pair adds in the same order as the upper one
    p = (threadIdx.x & off) ? padd(o, p) : padd(p, o);
  }
  return p;
}

// block-wide compensated sum of 5 pairs; the result is valid in thread 0
__device__ __forceinline__ void block_sum5(Pair *acc, double *lds) {
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    acc[n] = wave_sum(acc[n]);
    if (lane == 0) {
      lds[(wv * 5 + n) * 2] = acc[n].s;
      lds[(wv * 5 + n) * 2 + 1] = acc[n].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      Pair t{lds[n * 2], lds[n * 2 + 1]};
      for (int q = 1; q < nw; ++q) t = padd(t, Pair{lds[(q * 5 + n) * 2], lds[(q * 5 + n) * 2 + 1]});
      acc[n] = t;
    }
  __syncthreads();
}

__device__ __forceinline__ void sums_body(const DGrid &g, const DState &st, int lb, double *lds) {
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  size_t total = (size_t)6 * N * ni * N;
  Pair acc[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) acc[n] = Pair{0., 0.};
  if (t < total) {
    int j = g.js + (int)(t % N);
    size_t r = t / N;
    int i = g.is + (int)(r % ni);
    r /= ni;
    int k = g.ks + (int)(r % N), p = (int)(r / N);
    double vol = g.volume(p, k, j, i);
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (n < g.nv) acc[n].s = st.u[g.c(n, p, k, i, j)] * vol;
  }
  block_sum5(acc, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      st.part[((size_t)lb * 5 + n) * 2] = acc[n].s;
      st.part[((size_t)lb * 5 + n) * 2 + 1] = acc[n].c;
    }
}

// cell-shaped work list (one block per BLK interior cells)
__global__ void __launch_bounds__(BLK) k_sums(const Inst *__restrict__ I_, Work w) {
  __shared__ double lds[(BLK / 64) * 5 * 2];
  WORK_LOOP(w) {
    WORK_INST(w);
    sums_body(g, st, lb, lds);
  }
}

// one block per instance; cstart is the cell-shaped table (partials per
// instance), out[instance][5] receives the rounded compensated totals
__global__ void __launch_bounds__(BLK) k_sums_final(const Inst *__restrict__ I_, Work w, const int *__restrict__ cstart,
                                                   double *out) {
  __shared__ double lds[(BLK / 64) * 5 * 2];
  WORK_LOOP(w) {
    WORK_INST(w);
    (void)g;
    (void)lb;
    int nb = cstart[ii_ + 1] - cstart[ii_];
    Pair acc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[n] = Pair{0., 0.};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
      for (int n = 0; n < 5; ++n)
        acc[n] = padd(acc[n], Pair{st.part[((size_t)b * 5 + n) * 2],
                                   st.part[((size_t)b * 5 + n) * 2 + 1]});
    block_sum5(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
      for (int n = 0; n < 5; ++n) out[(size_t)ii_ * 5 + n] = acc[n].s + acc[n].c;
  }
}

// ===========================================================================
// Mixed precision (EXO_PREC_MIXED): FP32 WENO5 + Riemann on perturbations
// about a reference that is constant along the sweep, the reference-pressure
// part of the momentum flux removed (exo_mixed.hpp) and added back in FP64 by
// k_update<true>.  The decomposition scan, metric terms, sources, the
// implicit solve and the state stay FP64.

// x1 fluxes: FP64 decomposition (psf in LDS, stored for the update), FP32
// reconstruction of the perturbation and FP32 Riemann flux.
// FP64 series of the mixed form's decomposition (accurate to ~1e-15, far
// below the FP32 perturbations they feed).  atanh(t)/t for |t| <= 0.3, exp(x)
// for |x| <= 0.5, and ln((1 - exp(-l)) / l) from the Bernoulli-number series of
// ln((e^x - 1)/x) = x/2 + x^2/24 - x^4/2880 + x^6/181440 - x^8/9676800 + ...
__device__ __forceinline__ double atanh_over_t(double t2) {
  double s = 1.0 / 21;
#pragma unroll
  for (int m = 19; m >= 3; m -= 2) s = fma(s, t2, 1.0 / m);
  return fma(s, t2, 1.0);
}

__device__ __forceinline__ double exp_small(double x) {
  double s = 1.0 / 6402373705728000.0;  // 1/18!
  const double inv[17] = {1.0 / 355687428096000.0, 1.0 / 20922789888000.0,
                          1.0 / 1307674368000.0,   1.0 / 87178291200.0,
                          1.0 / 6227020800.0,      1.0 / 479001600.0,
                          1.0 / 39916800.0,        1.0 / 3628800.0,
                          1.0 / 362880.0,          1.0 / 40320.0,
                          1.0 / 5040.0,            1.0 / 720.0,
                          1.0 / 120.0,             1.0 / 24.0,
                          1.0 / 6.0,               0.5,
                          1.0};
#pragma unroll
  for (int m = 0; m < 17; ++m) s = fma(s, x, inv[m]);
  return fma(s, x, 1.0);
}

__device__ __forceinline__ double ln_expm1_ratio(double l) {  // ln((1 - e^-l) / l)
  double l2 = l * l;
  double s = fma(fma(fma(-1.0 / 9676800.0, l2, 1.0 / 181440.0), l2, -1.0 / 2880.0), l2,
                 1.0 / 24.0);
  return fma(s, l2, -0.5 * l);
}

// x1 fluxes of the mixed form: the FP64 face pressure psf as the reference
// computes it (psf in LDS and stored for the update), then the cell values
// psv and rho_sv and the face values rho_sf of the decomposition from one
// log-domain step per face: lambda = ln(a / b) = 2 atanh((a - b) / (a + b))
// for the faces a, b of a cell gives psv = (a - b) / lambda,
// rho_sf(b) = rho_sf(a) e^(-lambda / gamma) and
// rho_sv = rho_sf(a) ((1 - e^-lambda) / lambda)^(1/gamma).  Outside the
// series' range a cell falls back to the exact log / pow.
template <int W, int RK = -1>
__device__ __forceinline__ void flux_x1_mix_body(const DGrid &g, const DState &st, int lb,
                                                 double *lds) {
#pragma clang fp contract(fast)
  int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  int N = g.N;
  int nbx = (N * N + 63) / 64;
  int col = (lb % nbx) * 64 + lane;
  int p = lb / nbx;
  bool act = col < N * N;
  int k = g.ks + (act ? col / N : 0), j = g.js + (act ? col % N : 0);
  double *psf = lds;
#define PSF(ii) psf[(ii) * 64 + lane]
  int is = g.is, ie = g.ie;
  if (seg == 0 && act) psf_scan(g, st, p, k, j, psf, lane);
  __syncthreads();
  if (!act) return;
  int nf = ie + 2 - is;
  int f0 = is + (nf * seg) / W, f1 = is + (nf * (seg + 1)) / W;
  double ent0 = g.gamma;
  double ent1 = dlog(st.w[g.c(IPR, p, k, is, j)]) - ent0 * dlog(st.w[g.c(IDN, p, k, is, j)]);
  double ig = qdiv(1., ent0), e1 = dexp(qdiv(-ent1, ent0));
  // cell i from rho_sf at its lower face: psv, rho_sv and rho_sf at face i+1
  auto decomp = [&](int i, double rsf_lo, double *psv, double *rsv) -> double {
    double a = PSF(i), b = PSF(i + 1);
    double t = qdiv(a - b, a + b);
    if (fabs(t) > 0.3) {  // outside the series' range: the exact functions
      double ps = fabs(a - b) < 1.E-6 ? (a + b) / 2. : qdiv(a - b, dlog(qdiv(a, b)));
      *psv = ps;
      *rsv = dpow(ps, ig) * e1;
      return dpow(b, ig) * e1;
    }
    double lam = 2.0 * t * atanh_over_t(t * t);
    *psv = fabs(a - b) < 1.E-6 ? (a + b) / 2. : qdiv(a - b, lam);
    *rsv = rsf_lo * exp_small(ig * ln_expm1_ratio(lam));
    return rsf_lo * exp_small(-ig * lam);
  };
  float q[6][5];
  double rsf[4];  // rho_sf at the faces fi .. fi+3 of the window
  double rlo = dpow(PSF(f0 - 3), ig) * e1;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    int i = f0 - 3 + s;
    double psv, rsv;
    double rhi = decomp(i, rlo, &psv, &rsv);
    if (s >= 3) rsf[s - 3] = rlo;  // faces f0 .. f0+2
    rlo = rhi;
    q[s][IDN] = (float)(st.w[g.c(IDN, p, k, i, j)] - rsv);
    q[s][IVX] = (float)st.w[g.c(IVX, p, k, i, j)];
    q[s][IVY] = (float)st.w[g.c(IVY, p, k, i, j)];
    q[s][IVZ] = (float)st.w[g.c(IVZ, p, k, i, j)];
    q[s][IPR] = (float)(st.w[g.c(IPR, p, k, i, j)] - psv);
  }
  rsf[3] = rlo;  // face f0+3
  float cth = (float)g.T(T_CTH1, p, k, j), cosc = (float)g.T(T_COSC, p, k, j);
  float gamma = (float)g.gamma;
  for (int fi = f0; fi < f1; ++fi) {
    float wl[5], wr[5], fl[5];
    // the next window cell is loaded before this face is worked on
    bool more = fi + 1 < f1;
    double nxt[5];
    if (more)
#pragma unroll
      for (int n = 0; n < 5; ++n) nxt[n] = st.w[g.c(n, p, k, fi + 3, j)];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      wl[n] = weno5f(q[4][n], q[3][n], q[2][n], q[1][n], q[0][n]);
      wr[n] = weno5f(q[1][n], q[2][n], q[3][n], q[4][n], q[5][n]);
    }
    double pf = PSF(fi);
    float P = (float)pf;
    // RestoreFromPerturbation clamps (change_to_perturbation.cpp:174-189)
    if (wl[IPR] < -P) wl[IPR] = 0.f;
    if (wr[IPR] < -P) wr[IPR] = 0.f;
    float dsf = (float)rsf[0];
    wl[IDN] += dsf;
    if (wl[IDN] < 0.f) wl[IDN] = dsf;
    wr[IDN] += dsf;
    if (wr[IDN] < 0.f) wr[IDN] = dsf;
    riemann_pert(EOS_IDEAL, rs_of<RK>(g), IVX, gamma, wl, wr, P, fl, cth, 0.f, 1.f, cosc);
    for (int n = 0; n < 5; ++n) st.fflx[0][g.f(n, p, k, fi, j)] = fl[n];
    st.psf[g.f(0, p, k, fi, j)] = pf;
    if (more) {
#pragma unroll
      for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int n = 0; n < 5; ++n) q[s][n] = q[s + 1][n];
      int i = fi + 3;
      double psv, rsv;
      double rhi = decomp(i, rsf[3], &psv, &rsv);
      rsf[0] = rsf[1];
      rsf[1] = rsf[2];
      rsf[2] = rsf[3];
      rsf[3] = rhi;
      q[5][IDN] = (float)(nxt[IDN] - rsv);
      q[5][IVX] = (float)nxt[IVX];
      q[5][IVY] = (float)nxt[IVY];
      q[5][IVZ] = (float)nxt[IVZ];
      q[5][IPR] = (float)(nxt[IPR] - psv);
    }
  }
  // (a window sliding up the column, as in the FP64 sweep, measured 3-4 %
  // slower here: 0.091 -> 0.094 ms with two scalars on 6 x 64 x 64 x 64)
  for (int n = 0; n < g.nt; ++n)
    for (int fi = f0; fi < f1; ++fi) {
      float r[6];
      for (int s = 0; s < 6; ++s) r[s] = (float)st.r[g.c(n, p, k, fi - 3 + s, j)];
      float rl = weno5f(r[4], r[3], r[2], r[1], r[0]);
      float rr = weno5f(r[1], r[2], r[3], r[4], r[5]);
      float m = st.fflx[0][g.f(IDN, p, k, fi, j)];
      st.fsflx[0][g.f(n, p, k, fi, j)] = m >= 0.f ? m * rl : m * rr;
    }
#undef PSF
}

template <int W, int RK = -1>
__global__ void __launch_bounds__(64 * W) k_flux_x1_mix(const Inst *__restrict__ I_, Work w) {
  extern __shared__ double lds[];
  WORK_LOOP(w) {
    WORK_INST(w);
    flux_x1_mix_body<W, RK>(g, st, lb, lds);
    __syncthreads();  // psf is rebuilt by the next block of the loop
  }
}

// mixed ideal gas: the interior-side WENO5 state of a panel-edge face, moved
// into the neighbour's frame and slot
__device__ __forceinline__ void edge_mix_body(const DGrid &g, const DState &st, int lb,
                                              const ADDR(1) EdgeDesc *__restrict__ ed, int nedge) {
#pragma clang fp contract(fast)
  int t = lb * BLK + threadIdx.x;
  int ni = g.nx1;
  if (t >= nedge * ni) return;
  int e = t / ni, i = g.is + t % ni;
  EdgeDesc d = *(const EdgeDesc *)(ed + e);
  int p = d.src_panel, k = d.src_k, j = d.src_j;
  float q[5] = {0, 0, 0, 0, 0};
  // the FP32 perturbations k_prim wrote
  auto FW = [&](int n, int kk, int ii, int jj) -> float { return st.fw[g.c(n, p, kk, ii, jj)]; };
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (g.skipv(n)) continue;
    float v;
    if (d.src_dir == 0)
      v = weno5f(FW(n, k, i, j - 2), FW(n, k, i, j - 1),
                 FW(n, k, i, j), FW(n, k, i, j + 1),
                 FW(n, k, i, j + 2));
    else if (d.src_dir == 1)
      v = weno5f(FW(n, k, i, j + 1), FW(n, k, i, j),
                 FW(n, k, i, j - 1), FW(n, k, i, j - 2),
                 FW(n, k, i, j - 3));
    else if (d.src_dir == 2)
      v = weno5f(FW(n, k - 2, i, j), FW(n, k - 1, i, j),
                 FW(n, k, i, j), FW(n, k + 1, i, j),
                 FW(n, k + 2, i, j));
    else
      v = weno5f(FW(n, k + 1, i, j), FW(n, k, i, j),
                 FW(n, k - 1, i, j), FW(n, k - 2, i, j),
                 FW(n, k - 3, i, j));
    q[n] = v;
  }
  double U, V, vx, vy;
  vec_rll_from_abp(d.x2s, d.x3s, d.src_block, (double)q[IVY], (double)q[IVZ], &U, &V);
  vec_abp_from_rll(d.x2d, d.x3d, d.tgt_block, U, V, &vx, &vy);
  q[IVY] = (float)vx;
  q[IVZ] = (float)vy;
  size_t base = (((size_t)d.tgt_panel * 4 + d.tgt_dir) * g.N + d.tgt_a) * ni + (i - g.is);
  size_t vs = (size_t)6 * 4 * g.N * ni;
#pragma unroll
  for (int n = 0; n < 5; ++n)
    if (!g.skipv(n)) st.fedge[n * vs + base] = q[n];
}

__global__ void __launch_bounds__(BLK) k_edge_states_mix(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    edge_mix_body(g, st, lb, glb(I_[ii_].ed), I_[ii_].ned);
  }
}

// Shallow-water base flux of the mixed form: F(b) = (h u, h u^2 + h^2/2, h u v)
// for the primitive state b = (h, v2, v3) of a face's left cell, through the
// face's frame maps and covariant projection as riemann() applies them; with
// s = sinf and P = h^2/2 the maps reduce to
//   x2 (dir 1): (h s v2, h s v2^2 + P/s, h s v2 v3 - cosf P/s)
//   x3 (dir 2): (h s v3, h s v2 v3 - cosf P/s, h s v3^2 + P/s).
// f receives (IDN, -, IVY, IVZ).
template <int dir>
__device__ __forceinline__ void sw_base_flux(double h, double v2, double v3, double cf, double sf,
                                             double ct, double *f) {
  double hs = h * sf, Ps = qdiv(0.5 * h * h, sf);
  double fy, fz;
  if (dir == 1) {
    f[IDN] = hs * v2;
    fy = hs * v2 * v2 + Ps;
    fz = hs * v2 * v3 - cf * Ps;
  } else {
    f[IDN] = hs * v3;
    fy = hs * v2 * v3 - cf * Ps;
    fz = hs * v3 * v3 + Ps;
  }
  f[IVX] = 0.;
  f[IVY] = fy + fz * ct;
  f[IVZ] = fz + fy * ct;
}

template <int dir, bool SW, int RK = -1>
__device__ __forceinline__ void flux_h_mix_body(const DGrid &g, const DState &st, int lb) {
#pragma clang fp contract(fast)
  size_t t = (size_t)lb * BLK + threadIdx.x;
  int N = g.N, ni = g.nx1;
  int na = N + 1;
  size_t total = (size_t)6 * N * ni * na;
  if (t >= total) return;
  int kk, fa, i, p;
  if (dir == 1) {
    fa = (int)(t % na);
    size_t r = t / na;
    i = g.is + (int)(r % ni);
    r /= ni;
    kk = (int)(r % N);
    p = (int)(r / N);
  } else {
    // j fastest, then the face index k, then i: the six-row stencil along k
    // of consecutive faces is re-read within a few blocks, from L2
    kk = (int)(t % N);
    size_t r = t / N;
    fa = (int)(r % na);
    r /= na;
    i = g.is + (int)(r % ni);
    p = (int)(r / ni);
  }
  int k = dir == 1 ? g.ks + kk : g.ks + fa;
  int j = dir == 1 ? g.js + fa : g.js + kk;
  if (!SW) {
    float fl[5];
    face_flux_mix<dir, RK>(g, st, p, kk, fa, i, k, j, fl);
    float m = fl[IDN];
    ADDR(1) float *F = st.fflx[dir];
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (!g.skipv(n)) F[g.f(n, p, k, i, j)] = fl[n];
    for (int n = 0; n < g.nt; ++n) {
      float r[6];
      int nr = g.nv + n;  // the tracer's FP32 copy
      for (int s = 0; s < 6; ++s)
        r[s] = dir == 1 ? st.fw[g.c(nr, p, k, i, j - 3 + s)] : st.fw[g.c(nr, p, k - 3 + s, i, j)];
      float rl = weno5f(r[4], r[3], r[2], r[1], r[0]);
      float rr = weno5f(r[1], r[2], r[3], r[4], r[5]);
      st.fsflx[dir][g.f(n, p, k, i, j)] = m >= 0.f ? m * rl : m * rr;
    }
    return;
  }
  float wl[5] = {0, 0, 0, 0, 0}, wr[5] = {0, 0, 0, 0, 0}, fl[5];
  // shallow water: deltas about the face's left cell (riemann_sw_delta)
  constexpr bool sw = SW;
  double B[5] = {0., 0., 0., 0., 0.};
  if (sw) {
    int kL = dir == 1 ? k : k - 1, jL = dir == 1 ? j - 1 : j;
    B[IDN] = st.w[g.c(IDN, p, kL, i, jL)];
    B[IVY] = st.w[g.c(IVY, p, kL, i, jL)];
    B[IVZ] = st.w[g.c(IVZ, p, kL, i, jL)];
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    if (g.skipv(n)) continue;
    float v[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      int ks = dir == 1 ? k : k - 3 + s, js = dir == 1 ? j - 3 + s : j;
      v[s] = sw ? (float)(st.w[g.c(n, p, ks, i, js)] - B[n]) : st.fw[g.c(n, p, ks, i, js)];
    }
    wl[n] = weno5f(v[4], v[3], v[2], v[1], v[0]);
    wr[n] = weno5f(v[1], v[2], v[3], v[4], v[5]);
  }
  size_t vs = (size_t)6 * 4 * N * ni;
  int lo_side = dir == 1 ? 0 : 2;
  // panel edges: the neighbour's state (FP64 for shallow water, taken about B)
  if (fa == 0) {
    size_t b = (((size_t)p * 4 + lo_side) * N + kk) * ni + (i - g.is);
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (!g.skipv(n)) wl[n] = sw ? (float)(st.edge[n * vs + b] - B[n]) : st.fedge[n * vs + b];
  } else if (fa == N) {
    size_t b = (((size_t)p * 4 + lo_side + 1) * N + kk) * ni + (i - g.is);
#pragma unroll
    for (int n = 0; n < 5; ++n)
      if (!g.skipv(n)) wr[n] = sw ? (float)(st.edge[n * vs + b] - B[n]) : st.fedge[n * vs + b];
  }
  // full density for the ideal gas
  float P = 0.f;
  if (!sw) {
    float rref = (float)st.lref[2 * g.nc1 + i];
    wl[IDN] += rref;
    wr[IDN] += rref;
    P = (float)st.lref[g.nc1 + i];
  }
  double cosd, sind, cthd;
  if (dir == 1) {
    cosd = g.T(T_COS2, p, k, j);
    sind = g.T(T_SIN2, p, k, j);
    cthd = g.T(T_CTH2, p, k, j);
  } else {
    cosd = g.T(T_COS3, p, k, j);
    sind = g.T(T_SIN3, p, k, j);
    cthd = g.T(T_CTH3, p, k, j);
  }
  float cosf = (float)cosd, sinf = (float)sind, cth = (float)cthd;
  if (sw) {
    // FP64 face flux = base flux + FP32 delta; the panel-edge faces are
    // written by k_edge_flux in FP64
    float bf[5] = {(float)B[IDN], 0.f, (float)B[IVY], (float)B[IVZ], 0.f}, mb;
    riemann_sw_delta(rs_of<RK>(g), dir == 1 ? IVY : IVZ, bf, wl, wr, fl, cth, cosf, sinf, &mb);
    double fb[4];
    sw_base_flux<dir>(B[IDN], B[IVY], B[IVZ], cosd, sind, cthd, fb);
    ADDR(1) double *F = st.flx[dir];
    if (fa != 0 && fa != N) {
#pragma unroll
      for (int n = 0; n < 4; ++n)
        if (!g.skipv(n)) F[g.f(n, p, k, i, j)] = fb[n] + (double)fl[n];
    }
    double m = fb[IDN] + (double)fl[IDN];
    for (int n = 0; n < g.nt; ++n) {
      float r[6];
      for (int s = 0; s < 6; ++s)
        r[s] = (float)(dir == 1 ? st.r[g.c(n, p, k, i, j - 3 + s)] : st.r[g.c(n, p, k - 3 + s, i, j)]);
      float rl = weno5f(r[4], r[3], r[2], r[1], r[0]);
      float rr = weno5f(r[1], r[2], r[3], r[4], r[5]);
      st.sflx[dir][g.f(n, p, k, i, j)] = m >= 0.0 ? m * (double)rl : m * (double)rr;
    }
    return;
  }
}

// SW: the shallow-water form (FP64 face fluxes from FP64 base fluxes and
// FP32 deltas), a kernel of its own so the ideal-gas one compiles without it;
// RK: the Riemann solver of every instance of the launch, or -1
template <int dir, bool SW, int RK = -1>
__global__ void __launch_bounds__(BLK) k_flux_h_mix(const Inst *__restrict__ I_, Work w) {
  WORK_LOOP(w) {
    WORK_INST(w);
    flux_h_mix_body<dir, SW, RK>(g, st, lb);
  }
}

// single-face / scalar shims
__global__ void k_riemann_face(Grid g, int dir, int p, int k, int j, const double *wl0,
                               const double *wr0, double *flx) {
  double wl[5] = {0, 0, 0, 0, 0}, wr[5] = {0, 0, 0, 0, 0}, fl[5];
  for (int n = 0; n < g.nv; ++n) {
    wl[n] = wl0[n];
    wr[n] = wr0[n];
  }
  double cth, cosf = 0., sinf = 1., cosc = g.T(T_COSC, p, k, j);
  if (dir == 0) cth = g.T(T_CTH1, p, k, j);
  else if (dir == 1) { cth = g.T(T_CTH2, p, k, j); cosf = g.T(T_COS2, p, k, j); sinf = g.T(T_SIN2, p, k, j); }
  else { cth = g.T(T_CTH3, p, k, j); cosf = g.T(T_COS3, p, k, j); sinf = g.T(T_SIN3, p, k, j); }
  riemann<double>(g.eos, g.rs, dir == 0 ? IVX : (dir == 1 ? IVY : IVZ), g.gamma, wl, wr, fl,
                  cth, cosf, sinf, cosc);
  for (int n = 0; n < g.nv; ++n) flx[n] = fl[n];
}

__global__ void k_weno_batch(const double *in5, double *out, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double *a = in5 + 5 * (size_t)t;
  out[t] = weno5(a[0], a[1], a[2], a[3], a[4]);
}

// op 0 exp, 1 log, 2 pow(x, y), 3 div_rn(x[2t], x[2t+1], 1 / x[2t+1])
__global__ void k_math_batch(int op, const double *x, double y, double *out, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (op == 3 || op == 4) {
    double a = x[2 * t], d = x[2 * t + 1];
    out[t] = op == 3 ? div_rn(a, d, 1.0 / d) : qdiv(a, d);
    return;
  }
  out[t] = op == 0 ? dexp(x[t]) : (op == 1 ? dlog(x[t]) : dpow(x[t], y));
}

// ---------------------------------------------------------------------------
// output fields in the per-block file layout (src/outputs/netcdf.cpp:439-447:
// float, x1 outermost, x3 fastest)
constexpr int OUT_MAX = 24;
enum { OP_COPY = 0, OP_TAB, OP_U, OP_V };
struct OutMap {
  const double *src[OUT_MAX];  // OP_COPY: the field's cell array; OP_TAB: a table plane
  int op[OUT_MAX];
  int flip[OUT_MAX];           // x1 vector component: negated in the x1 mirror
};

// out[p][q][i'][j'][k'] from the device layout [k][i][j]: one workgroup moves a
// 64 x 64 (k, j) tile of one field at one x1 level through LDS, reading rows
// of 64 doubles along x2 and writing rows of 64 floats along x3.
//   * The lat-lon wind takes the contravariant (v2, v3) of the cell and its
//     column's factors in the grouping of CubedSphere::GetUV
//     (cubed_sphere.cpp:132-145).
//   * x1 ghost levels are the reflecting mirrors of the interior (the x1
//     component negated), as the reflecting boundary leaves w.
//   * Ghost cells outside both the x2 and the x3 interior range (the panel
//     corners, which no stencil reads and no exchange fills here) are NaN,
//     except for the geometric fields.
__global__ void __launch_bounds__(256) k_out_fields(Grid g, OutMap m, const double *wy,
                                                    const double *wz, const double *uvt,
                                                    int nq, int ghost, float *out) {
  __shared__ float tile[64][65];
  int G = ghost ? NG : 0;
  int n1 = g.nx1 + 2 * G, n2 = g.N + 2 * G, n3 = n2;
  int ntj = (n2 + 63) >> 6;
  int tj = blockIdx.x % ntj, tk = blockIdx.x / ntj;
  int i = blockIdx.y, p = blockIdx.z / nq, q = blockIdx.z % nq;
  // the interior level an x1 ghost level mirrors: reflections repeated until
  // inside (more than one only when nx1 < 3), the x1 component negated once
  // per reflection
  int isrc = g.is - G + i;
  bool mir = false;
  for (int r = 0; r < 4 && (isrc < g.is || isrc > g.ie); ++r) {
    isrc = isrc < g.is ? 2 * g.is - 1 - isrc : 2 * g.ie + 1 - isrc;
    mir = !mir;
  }
  int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  int op = m.op[q];
  size_t tp = (size_t)6 * g.nc3 * g.nc2;
  for (int r = ly; r < 64; r += 4) {
    int kk = tk * 64 + r, jj = tj * 64 + lx;
    if (kk < n3 && jj < n2) {
      int k = g.ks - G + kk, j = g.js - G + jj;
      bool corner = (k < g.ks || k > g.ke) && (j < g.js || j > g.je);
      size_t c = g.c(0, p, k, isrc, j);
      size_t t = ((size_t)p * g.nc3 + k) * g.nc2 + j;
      double v;
      if (op == OP_TAB) {
        v = m.src[q][t];
      } else if (corner) {
        v = __builtin_nan("");
      } else if (op == OP_COPY) {
        v = m.src[q][c];
        if (mir && m.flip[q]) v = -v;
      } else {
        double V2 = wy[c], V3 = wz[c];
        double f1 = uvt[3 * tp + t], f2 = uvt[4 * tp + t];
        double f3 = uvt[5 * tp + t], f4 = uvt[6 * tp + t];
        if (p == 0) v = op == OP_U ? -(f1 * V2 + f2 * V3) : f3 * V2 - f4 * V3;
        else if (p == 4) v = op == OP_U ? -(f1 * V2 - f2 * V3) : f3 * V2 + f4 * V3;
        else v = op == OP_U ? -(f1 * V3) : f2 * V3 - V2;
      }
      tile[r][lx] = (float)v;
    }
  }
  __syncthreads();
  for (int r = ly; r < 64; r += 4) {
    int jj = tj * 64 + r, kk = tk * 64 + lx;
    if (kk < n3 && jj < n2)
      out[((((size_t)p * nq + q) * n1 + i) * n2 + jj) * n3 + kk] = tile[lx][r];
  }
}

}  // namespace

// ===========================================================================
// launches: every kernel runs over a work list of instances (one for a
// handle, many for a batch)
namespace {

// launch kinds, each with its own block-prefix table over the instances
enum {
  K_PRIM = 0, K_GHOST_U, K_GHOST_S, K_X1, K_X1_MIX, K_EDGE, K_EDGE_MIX, K_FACE, K_FACE_MIX,
  K_UPD, K_UPD_MIX, K_IMPL, K_ADD_DU, K_CELLS, K_ONE, K_DT, K_EFLX, K_FACE_MIX_SW,
  // within a step: the full primitive conversion for the ideal gas, the ghost
  // strips for shallow water (its stage update writes the interior)
  K_PRIM_REST, K_PRIM_GHOST,
  // the partial (3x3) column correction
  K_IMPL3,
  // within a stage: the full 5x5 solve, the shallow-water update, the partial
  // 3x3 solve and u += du without a solve (implicit_flag 0 or 8); the column
  // steps form the stage's weighted average of u (the ideal gas's update,
  // k_face2_update below, forms only the tendency)
  K_IMPL_AVG, K_UPD_RS, K_IMPL3_AVG, K_ADD_DU_AVG,
  // within a stage, the FP64 ideal gas takes its x2 faces and update in one
  // kernel (k_face2_update, without and with scalars), FP64 shallow water its
  // separate x2 face kernel (the mixed forms have their own kinds)
  K_FUSE2, K_FACE_NF, K_FUSE2_S, K_FUSE2_MIX, K_FUSE2_MIX_S, K_NKIND
};

// kernel ids of the profile (exo_profile_end)
enum { KID_PRIM = 0, KID_GHOST, KID_FLUX_X1, KID_EDGE, KID_FLUX_X2, KID_FLUX_X3, KID_UPDATE,
       KID_IMPLICIT, KID_DT, KID_OUTPUT, KID_COUNT };

unsigned grid1(size_t n, unsigned b = BLK) { return (unsigned)((n + b - 1) / b); }

// density and scalar mixing-ratio floors, sqrt(1024 FLT_MIN)
const double DFLOOR = std::sqrt(1024 * (double)FLT_MIN);
const double SFLOOR = std::sqrt(1024 * (double)FLT_MIN);

struct Launcher {
  hipStream_t stream = nullptr;
  int ninst = 0;
  Inst *d_inst = nullptr;
  int *d_bstart = nullptr;  // [K_NKIND][ninst + 1]
  int nblk[K_NKIND] = {};
  int cap[K_NKIND] = {};    // workgroups of a persistent launch (0: one per block)
  int shape[K_NKIND] = {};  // tuned block shape of a kind (0: the default): the x1
                            // kinds' waves per 64-column block (2, 4 or 8)
  int eos = -1;             // equation of state of every instance (-1: they differ)
  int rs = -1;              // Riemann solver of every instance (-1: they differ)
  std::vector<Inst> h_inst;     // host copies of the instances (the tuner reads them)
  size_t lds_x1 = 0;        // dynamic LDS of the x1 kernels (largest instance)
  // live per-kernel timing with HIP events on the stream
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_kind;

  Work work(int kind) const {
    return Work{d_bstart + (size_t)kind * (ninst + 1), ninst, nblk[kind]};
  }
  unsigned grid(int kind) const {
    int n = nblk[kind];
    return (unsigned)(cap[kind] > 0 && cap[kind] < n ? cap[kind] : n);
  }
};

}  // namespace

struct exo_handle {
  Setup setup;
  Grid g;
  State st;
  Launcher L;                   // this instance alone
  hipStream_t stream = nullptr; // == L.stream
  std::string err;
  std::vector<void *> allocs;
  GhostDesc *d_ghost = nullptr;
  EdgeDesc *d_edge = nullptr;
  double *d_x1 = nullptr, *d_tab = nullptr;
  double *d_sums = nullptr;     // [5] totals of exo_conserved_sums
  double *d_uov = nullptr;      // per-column factors of the "uov" output (first use)
  float *d_out = nullptr;       // output fields (grown on demand)
  size_t out_cap = 0;           // floats in d_out
  double *d_snap = nullptr;     // Mesh::SaveAllStates snapshot: u, w, s, r, dtv (first use)
  int redo_attempts = 0;        // redone steps over the handle's life (main.cpp.in:127)
  bool have_state = false;
  bool mixed = false;
  // one SSPRK3 step captured as a HIP graph and replayed by exo_step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt_fixed = 0.0;

  void *dalloc(size_t bytes) {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, bytes));
    HIPCHK(hipMemsetAsync(p, 0, bytes, stream));
    allocs.push_back(p);
    return p;
  }
  size_t cells_bytes(int nvar) const { return (size_t)nvar * g.vstride * sizeof(double); }
  size_t face_bytes(int nvar) const { return (size_t)nvar * g.fvstride * sizeof(double); }
};

// A set of instances advanced together: one launch per phase over all of
// them, optionally persistent, each step replayed from a captured HIP graph.
struct exo_batch {
  std::vector<exo_handle *> members;
  Launcher L;
  std::string err;
  int use_graph = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt_fixed = 0.0;
  double *d_sums = nullptr;  // [members][5]
};

namespace {

hipEvent_t next_event(Launcher &L) {
  if (L.ev_used == L.ev_pool.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    L.ev_pool.push_back(e);
  }
  return L.ev_pool[L.ev_used++];
}

// blocks of one instance in a launch of the given kind (0: not applicable)
int kind_blocks(int kind, const Grid &g, bool mixed, int ngd, int ned) {
  bool ideal = g.eos == EOS_IDEAL;
  size_t cells = (size_t)6 * g.N * g.N * g.nx1;
  size_t faces = (size_t)6 * g.N * g.nx1 * (g.N + 1);
  size_t cols = (size_t)g.N * g.N;
  // the full 5x5 column solve (implicit_flag bits 1 and 8)
  bool full = ideal && (g.implicit & 9) == 9;
  switch (kind) {
    case K_PRIM: return (int)grid1((size_t)6 * g.nc3 * g.nx1 * g.nc2);
    case K_PRIM_REST: return ideal ? (int)grid1((size_t)6 * g.nc3 * g.nx1 * g.nc2) : 0;
    case K_PRIM_GHOST: return !ideal ? (int)grid1((size_t)6 * 4 * NG * g.N * g.nx1) : 0;
    case K_GHOST_U: return (int)grid1((size_t)ngd * g.nx1);
    case K_GHOST_S: return g.nt ? (int)grid1((size_t)ngd * g.nx1) : 0;
    case K_X1: return ideal && !mixed ? (int)grid1(cols, 64) * 6 : 0;
    case K_X1_MIX: return ideal && mixed ? (int)grid1(cols, 64) * 6 : 0;
    // mixed shallow water takes FP64 edge states (its faces are taken about an
    // FP64 base state)
    case K_EDGE: return !mixed || !ideal ? (int)grid1((size_t)ned * g.nx1) : 0;
    case K_EDGE_MIX: return mixed && ideal ? (int)grid1((size_t)ned * g.nx1) : 0;
    case K_FACE: return !mixed ? (int)grid1(faces) : 0;
    case K_FACE_MIX: return mixed && ideal ? (int)grid1(faces) : 0;
    case K_FACE_MIX_SW: return mixed && !ideal ? (int)grid1(faces) : 0;
    // mixed shallow water stores FP64 face fluxes: the FP64 update
    case K_UPD: return !mixed || !ideal ? (int)grid1(cells) : 0;
    case K_UPD_MIX: return mixed && ideal ? (int)grid1(cells) : 0;
    case K_IMPL_AVG: return full ? (int)grid1(cols, 64) * 6 : 0;
    case K_UPD_RS: return !ideal ? (int)grid1(cells) : 0;
    case K_IMPL3_AVG:
      return ideal && (g.implicit & 1) && !(g.implicit & 8) ? (int)grid1(cols, 64) * 6 : 0;
    case K_ADD_DU_AVG: return ideal && !(g.implicit & 1) ? (int)grid1(cells) : 0;
    case K_FUSE2:
    case K_FUSE2_S:
    case K_FUSE2_MIX:
    case K_FUSE2_MIX_S: {
      bool mk = kind == K_FUSE2_MIX || kind == K_FUSE2_MIX_S;
      bool sk = kind == K_FUSE2_S || kind == K_FUSE2_MIX_S;
      if (!ideal || mixed != mk || sk != (g.nt > 0)) return 0;
      size_t waves = (faces + F2_FACES - 1) / F2_FACES;
      return (int)((waves + BLK / 64 - 1) / (BLK / 64));
    }
    case K_FACE_NF: return !mixed && !ideal ? (int)grid1(faces) : 0;
    // implicit_flag bit 1 solves the x1 columns, bit 8 with the full 5x5
    // system (solve_implicit_3d.cpp:102-135); without bit 1, u += du
    case K_IMPL: return ideal && (g.implicit & 1) && (g.implicit & 8) ? (int)grid1(cols, 64) * 6 : 0;
    case K_IMPL3:
      return ideal && (g.implicit & 1) && !(g.implicit & 8) ? (int)grid1(cols, 64) * 6 : 0;
    case K_ADD_DU: return ideal && !(g.implicit & 1) ? (int)grid1(cells) : 0;
    case K_CELLS: return (int)grid1(cells);
    case K_ONE: return 1;
    case K_DT: return (int)grid1(cells, BLK * DTC);
    case K_EFLX: return mixed && !ideal ? (int)grid1((size_t)6 * 4 * g.N * g.nx1) : 0;
  }
  return 0;
}

// uploads the instance list and block tables of the given handles
void launcher_setup(Launcher &L, const std::vector<exo_handle *> &hs,
                    std::vector<void *> &allocs) {
  int n = (int)hs.size();
  L.ninst = n;
  std::vector<Inst> inst(n);
  std::vector<int> tab((size_t)K_NKIND * (n + 1));
  L.lds_x1 = 0;
  for (int i = 0; i < n; ++i) {
    const exo_handle *h = hs[i];
    inst[i].g = h->g;
    inst[i].st = h->st;
    inst[i].gd = h->d_ghost;
    inst[i].ed = h->d_edge;
    inst[i].ngd = (int)h->setup.ghosts.size();
    inst[i].ned = (int)h->setup.edges.size();
    L.lds_x1 = std::max(L.lds_x1, (size_t)(h->g.nc1 + 1) * 64 * sizeof(double));
  }
  L.eos = n > 0 ? hs[0]->g.eos : -1;
  L.rs = n > 0 ? hs[0]->g.rs : -1;
  for (int i = 1; i < n; ++i) {
    if (hs[i]->g.eos != L.eos) L.eos = -1;
    if (hs[i]->g.rs != L.rs) L.rs = -1;
  }
  L.h_inst = inst;
  for (int kind = 0; kind < K_NKIND; ++kind) {
    long long acc = 0;
    for (int i = 0; i < n; ++i) {
      tab[(size_t)kind * (n + 1) + i] = (int)acc;
      acc += kind_blocks(kind, hs[i]->g, hs[i]->mixed, inst[i].ngd, inst[i].ned);
    }
    if (acc > 0x7fffffffLL) throw std::string("work list too large for one launch");
    tab[(size_t)kind * (n + 1) + n] = (int)acc;
    L.nblk[kind] = (int)acc;
  }
  {
    int dev = 0, maxlds = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    bool need_x1 = L.nblk[K_X1] > 0 || L.nblk[K_X1_MIX] > 0;
    for (const exo_handle *h : hs)
      if ((size_t)6 * h->g.nc3 * h->g.nc2 * h->g.nc1 >= ((size_t)1 << 32))
        throw std::string("instance too large: 2^32 cells or more");
    for (const exo_handle *h : hs)
      if (h->g.nc1 > IMPL_MAXNC1 && h->g.eos == EXO_EOS_IDEAL && (h->g.implicit & 1))
        throw std::string("nx1 too large for the implicit column solve (nc1 <= ") +
            std::to_string(IMPL_MAXNC1) + ")";
    if (need_x1 && L.lds_x1 > (size_t)maxlds)
      throw std::string("nx1 too large: the x1 column kernel needs ") +
          std::to_string(L.lds_x1) + " bytes of LDS, the device allows " +
          std::to_string(maxlds);
  }
  void *p;
  HIPCHK(hipMalloc(&p, sizeof(Inst) * n));
  allocs.push_back(p);
  L.d_inst = (Inst *)p;
  HIPCHK(hipMalloc(&p, sizeof(int) * tab.size()));
  allocs.push_back(p);
  L.d_bstart = (int *)p;
  HIPCHK(hipMemcpyAsync(L.d_inst, inst.data(), sizeof(Inst) * n, hipMemcpyHostToDevice, L.stream));
  HIPCHK(hipMemcpyAsync(L.d_bstart, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice,
                        L.stream));
  HIPCHK(hipStreamSynchronize(L.stream));
}

// a kind's block shape (the x1 sweeps' waves per block; the block count of a
// launch does not depend on it)
void set_shape(Launcher &L, int kind, int shape) {
  bool x1 = kind == K_X1 || kind == K_X1_MIX;
  bool ok = shape == 0 || (x1 && (shape == 2 || shape == 4 || shape == 8));
  if (!ok) throw std::string("no block shape ") + std::to_string(shape) + " for launch kind " +
               std::to_string(kind);
  L.shape[kind] = shape;
}

// one launch of a work-list kernel; kid < 0: not timed
template <typename... P, typename... A>
void launch(Launcher &L, int kind, int kid, void (*kern)(const Inst *, Work, P...), unsigned block,
            size_t lds, A... args) {
  if (L.nblk[kind] == 0) return;
  bool t = L.prof && kid >= 0;
  if (t) {
    L.ev_kind.push_back(kid);
    HIPCHK(hipEventRecord(next_event(L), L.stream));
  }
  kern<<<dim3(L.grid(kind)), dim3(block), lds, L.stream>>>(L.d_inst, L.work(kind), args...);
  if (t) HIPCHK(hipEventRecord(next_event(L), L.stream));
}

void make_grid(const Setup &s, Grid &g) {
  g.N = s.N;
  g.nx1 = s.nx1;
  g.nc1 = s.nc1;
  g.nc2 = s.nc2;
  g.nc3 = s.nc3;
  g.jp = (s.nc2 + 1 + 7) / 8 * 8;
  g.nv = s.nv;
  g.nt = s.nt;
  g.is = s.is; g.ie = s.ie; g.js = s.js; g.je = s.je; g.ks = s.ks; g.ke = s.ke;
  g.eos = s.cfg.eos;
  g.rs = s.cfg.rsolver;
  g.coriolis = s.cfg.coriolis;
  g.implicit = s.cfg.implicit_flag;
  g.cfl_x1 = s.cfg.cfl_x1;
  g.gamma = s.cfg.gamma;
  g.grav = s.cfg.eos == EXO_EOS_IDEAL ? s.cfg.grav : 0.;
  g.cfl = s.cfg.cfl;
  g.plane = (size_t)g.nc1 * g.jp;
  g.pstride = (size_t)g.nc3 * g.plane;
  g.vstride = 6 * g.pstride;
  g.fplane = (size_t)(g.nc1 + 1) * g.jp;
  g.fpstride = (size_t)(g.nc3 + 1) * g.fplane;
  g.fvstride = 6 * g.fpstride;
  g.t2stride = (size_t)6 * g.nc3 * g.nc2;
}

// Device bytes one instance allocates (the list init_handle allocates, each
// rounded up to the allocation granule), for admission by memory footprint.
size_t footprint_bytes(const Setup &s) {
  Grid g;
  make_grid(s, g);
  const size_t gran = (size_t)64 << 10;
  size_t tot = 0;
  auto add = [&](size_t b) { tot += (b + gran - 1) / gran * gran; };
  bool mixed = s.cfg.precision == EXO_PREC_MIXED;
  size_t cell = g.vstride * sizeof(double), face = g.fvstride * sizeof(double);
  size_t edge = (size_t)6 * 4 * g.N * g.nx1;
  add(3 * (s.nc1 + 1) * sizeof(double));
  add((s.tab2.size() + s.tab1.size()) * sizeof(double));
  add(s.ghosts.size() * sizeof(GhostDesc));
  add(s.edges.size() * sizeof(EdgeDesc));
  for (int q = 0; q < 4; ++q) add(cell * g.nv);           // u, w, u1, du
  if (g.nt)
    for (int q = 0; q < 3; ++q) add(cell * g.nt);         // s, r, s1
  // FP32 face fluxes and edge states for the mixed ideal gas only
  bool f32 = mixed && g.eos == EXO_EOS_IDEAL;
  if (!f32) {
    for (int d = 0; d < 3; ++d) add(face * g.nv);
    if (g.nt)
      for (int d = 0; d < 3; ++d) add(face * g.nt);
    add(edge * g.nv * sizeof(double));
  }
  if (g.eos == EXO_EOS_IDEAL && (g.implicit & 1)) add(cell * (g.implicit & 8 ? 30 : 12));
  if (f32) {
    for (int d = 0; d < 3; ++d) {
      add(g.fvstride * sizeof(float) * g.nv);
      if (g.nt) add(g.fvstride * sizeof(float) * g.nt);
    }
    add(edge * g.nv * sizeof(float));
    add(face);
    add(g.vstride * sizeof(float) * (g.nv + g.nt));  // fw
  }
  if (mixed) add(3 * (size_t)g.nc1 * sizeof(double));
  add(sizeof(unsigned long long));
  add(2 * sizeof(double));
  add(sizeof(int));
  add((size_t)grid1((size_t)6 * g.N * g.N * g.nx1) * 10 * sizeof(double));
  add(5 * sizeof(double));
  add(sizeof(Inst));
  add(sizeof(int) * K_NKIND * 2);
  return tot;
}

void init_handle(exo_handle *h) {
  Setup &s = h->setup;
  Grid &g = h->g;
  make_grid(s, g);

  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->L.stream = h->stream;
  // geometry
  std::vector<double> x1(3 * (s.nc1 + 1), 0.);
  for (int i = 0; i <= s.nc1; ++i) x1[i] = s.x1f[i];
  for (int i = 0; i < s.nc1; ++i) x1[s.nc1 + 1 + i] = s.x1v[i];
  for (int i = 0; i < s.nc1; ++i) x1[2 * (s.nc1 + 1) + i] = s.dx1f[i];
  h->d_x1 = (double *)h->dalloc(x1.size() * sizeof(double));
  HIPCHK(hipMemcpyAsync(h->d_x1, x1.data(), x1.size() * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  g.x1f = h->d_x1;
  g.x1v = h->d_x1 + (s.nc1 + 1);
  g.dx1f = h->d_x1 + 2 * (s.nc1 + 1);
  // the 2-D tables, then the per-row ones
  h->d_tab = (double *)h->dalloc((s.tab2.size() + s.tab1.size()) * sizeof(double));
  HIPCHK(hipMemcpyAsync(h->d_tab, s.tab2.data(), s.tab2.size() * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->d_tab + s.tab2.size(), s.tab1.data(), s.tab1.size() * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
  g.tab = h->d_tab;
  g.tab1 = h->d_tab + s.tab2.size();
  h->d_ghost = (GhostDesc *)h->dalloc(s.ghosts.size() * sizeof(GhostDesc));
  HIPCHK(hipMemcpyAsync(h->d_ghost, s.ghosts.data(), s.ghosts.size() * sizeof(GhostDesc),
                        hipMemcpyHostToDevice, h->stream));
  h->d_edge = (EdgeDesc *)h->dalloc(s.edges.size() * sizeof(EdgeDesc));
  HIPCHK(hipMemcpyAsync(h->d_edge, s.edges.data(), s.edges.size() * sizeof(EdgeDesc),
                        hipMemcpyHostToDevice, h->stream));
  // state
  State &st = h->st;
  std::memset(&st, 0, sizeof(st));
  st.u = (double *)h->dalloc(h->cells_bytes(g.nv));
  st.w = (double *)h->dalloc(h->cells_bytes(g.nv));
  st.u1 = (double *)h->dalloc(h->cells_bytes(g.nv));
  st.du = (double *)h->dalloc(h->cells_bytes(g.nv));
  if (g.nt) {
    st.s = (double *)h->dalloc(h->cells_bytes(g.nt));
    st.r = (double *)h->dalloc(h->cells_bytes(g.nt));
    st.s1 = (double *)h->dalloc(h->cells_bytes(g.nt));
  }
  h->mixed = s.cfg.precision == EXO_PREC_MIXED;
  // FP32 face fluxes and edge states for the mixed ideal gas only: mixed
  // shallow water forms FP64 face fluxes from FP64 base fluxes and FP32
  // deltas (flux_h_mix_body), the rest of its stage is the FP64 one
  bool f32 = h->mixed && g.eos == EXO_EOS_IDEAL;
  if (!f32) {
    for (int d = 0; d < 3; ++d) st.flx[d] = (double *)h->dalloc(h->face_bytes(g.nv));
    if (g.nt)
      for (int d = 0; d < 3; ++d) st.sflx[d] = (double *)h->dalloc(h->face_bytes(g.nt));
    st.edge = (double *)h->dalloc((size_t)g.nv * 6 * 4 * g.N * g.nx1 * sizeof(double));
  }
  if (g.eos == EXO_EOS_IDEAL && (g.implicit & 1)) {
    // the column solve's [block][level][pair][64 lanes] scratch: 15 pairs for
    // the 5x5 system (k_implicit), 6 for the 3x3 (k_implicit3)
    int per = g.implicit & 8 ? 30 : 12;
    size_t need = (size_t)6 * ((g.N * g.N + 63) / 64) * g.nx1 * per * 64;
    if (need > (size_t)per * g.vstride) throw std::string("implicit scratch layout does not fit");
    st.coef = (double *)h->dalloc(h->cells_bytes(per));
  }
  if (f32) {
    size_t fb = (size_t)g.fvstride * sizeof(float);
    for (int d = 0; d < 3; ++d) {
      st.fflx[d] = (float *)h->dalloc(fb * g.nv);
      if (g.nt) st.fsflx[d] = (float *)h->dalloc(fb * g.nt);
    }
    st.fedge = (float *)h->dalloc((size_t)g.nv * 6 * 4 * g.N * g.nx1 * sizeof(float));
    st.psf = (double *)h->dalloc(g.fvstride * sizeof(double));
    st.fw = (float *)h->dalloc((size_t)(g.nv + g.nt) * g.vstride * sizeof(float));
  }
  if (h->mixed) st.lref = (double *)h->dalloc((size_t)3 * g.nc1 * sizeof(double));
  st.dtbits = (unsigned long long *)h->dalloc(sizeof(unsigned long long));
  st.dtv = (double *)h->dalloc(2 * sizeof(double));
  st.bad = (int *)h->dalloc(sizeof(int));
  st.part = (double *)h->dalloc((size_t)grid1((size_t)6 * g.N * g.N * g.nx1) * 10 * sizeof(double));
  h->d_sums = (double *)h->dalloc(5 * sizeof(double));
  double dt0[2] = {0., DBL_MAX * 0.4};
  HIPCHK(hipMemcpyAsync(st.dtv, dt0, sizeof(dt0), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  launcher_setup(h->L, std::vector<exo_handle *>{h}, h->allocs);
}

// the full primitive conversion specialised for the launch's equation of state
void launch_prim(Launcher &L, int kid) {
  if (L.eos == EOS_SW) launch(L, K_PRIM, kid, k_prim<EOS_SW>, BLK, 0, DFLOOR, SFLOOR);
  else if (L.eos == EOS_IDEAL) launch(L, K_PRIM, kid, k_prim<EOS_IDEAL>, BLK, 0, DFLOOR, SFLOOR);
  else launch(L, K_PRIM, kid, k_prim<>, BLK, 0, DFLOOR, SFLOOR);
}

void launch_fill_ghosts(Launcher &L) {
  launch(L, K_GHOST_U, KID_GHOST, k_ghost, BLK, 0, 0);
  launch(L, K_GHOST_S, KID_GHOST, k_ghost, BLK, 0, 1);
  launch_prim(L, KID_PRIM);
  HIPCHK(hipGetLastError());
}

// Calls f with the launch's Riemann solver as a compile-time constant when
// all its instances share one of the two that equation of state E allows,
// else with -1 (each instance's own, read in the kernel).
template <int E, typename F>
void with_solver(const Launcher &L, F &&f) {
  constexpr int A = E == EOS_SW ? RS_ROE_SW : RS_HLLC;
  constexpr int B = E == EOS_SW ? RS_LMARS_SW : RS_LMARS;
  if (L.rs == A) f(std::integral_constant<int, A>{});
  else if (L.rs == B) f(std::integral_constant<int, B>{});
  else f(std::integral_constant<int, -1>{});
}

// the x1 sweep of one kind (ideal gas) with its tuned waves per block; the
// mixed one specialised for the launch's Riemann solver (the FP64 one is
// faster without: 0.237 against 0.277 ms with LMARS on 6 x 96 x 96 x 64)
template <int W>
void launch_x1_w(Launcher &L, int kind, int kid) {
  if (kind == K_X1_MIX)
    with_solver<EOS_IDEAL>(L, [&](auto R) {
      launch(L, kind, kid, k_flux_x1_mix<W, decltype(R)::value>, 64 * W, L.lds_x1);
    });
  else launch(L, kind, kid, k_flux_x1<W>, 64 * W, L.lds_x1);
}

void launch_x1(Launcher &L, int kind, int kid) {
  switch (L.shape[kind] ? L.shape[kind] : X1W) {
    case 2: launch_x1_w<2>(L, kind, kid); break;
    case 8: launch_x1_w<8>(L, kind, kid); break;
    default: launch_x1_w<4>(L, kind, kid);
  }
}

// the FP64 face kernel specialised for the launch's equation of state and
// Riemann solver
template <int dir>
void launch_face64(Launcher &L, int kind, int kid) {
  if (L.eos == EOS_SW)
    with_solver<EOS_SW>(L, [&](auto R) {
      launch(L, kind, kid, k_flux_h<dir, EOS_SW, decltype(R)::value>, BLK, 0);
    });
  else if (L.eos == EOS_IDEAL)
    with_solver<EOS_IDEAL>(L, [&](auto R) {
      launch(L, kind, kid, k_flux_h<dir, EOS_IDEAL, decltype(R)::value>, BLK, 0);
    });
  else launch(L, kind, kid, k_flux_h<dir>, BLK, 0);
}

// the mixed face kernel: the ideal gas's specialised for the launch's Riemann
// solver; the shallow-water form (SW) is faster without (1.03 against 1.10 ms
// per step with Roe on the 256-instance ensemble)
template <int dir, bool SW>
void launch_face_mix(Launcher &L, int kind, int kid) {
  if constexpr (SW)
    launch(L, kind, kid, k_flux_h_mix<dir, true>, BLK, 0);
  else
    with_solver<EOS_IDEAL>(L, [&](auto R) {
      launch(L, kind, kid, k_flux_h_mix<dir, false, decltype(R)::value>, BLK, 0);
    });
}

// fused: within a stage, the x2 faces of the instances k_face2_update covers
// are left to it
void launch_fluxes(Launcher &L, bool fused = false) {
  launch_x1(L, K_X1, KID_FLUX_X1);
  launch_x1(L, K_X1_MIX, KID_FLUX_X1);
  launch(L, K_EDGE, KID_EDGE, k_edge_states, BLK, 0);
  launch(L, K_EDGE_MIX, KID_EDGE, k_edge_states_mix, BLK, 0);
  launch(L, K_EFLX, KID_EDGE, k_edge_flux, BLK, 0);
  launch_face64<1>(L, fused ? K_FACE_NF : K_FACE, KID_FLUX_X2);
  launch_face64<2>(L, K_FACE, KID_FLUX_X3);
  if (!fused) launch_face_mix<1, false>(L, K_FACE_MIX, KID_FLUX_X2);
  launch_face_mix<2, false>(L, K_FACE_MIX, KID_FLUX_X3);
  launch_face_mix<1, true>(L, K_FACE_MIX_SW, KID_FLUX_X2);
  launch_face_mix<2, true>(L, K_FACE_MIX_SW, KID_FLUX_X3);
  HIPCHK(hipGetLastError());
}

const double SS_DELTA[3] = {1.0, 0.0, 0.0};
const double SS_G1[3] = {0.0, 0.25, 2.0 / 3.0};
const double SS_G2[3] = {1.0, 0.75, 1.0 / 3.0};
const double SS_BETA[3] = {1.0, 0.25, 2.0 / 3.0};

// the fused x2-face update (ideal gas) with the launch's Riemann solver; the
// mixed form with scalars is faster without (0.139 against 0.174 ms with LMARS
// and two scalars on 6 x 64 x 64 x 64)
template <bool SCAL, bool MIX>
void launch_fuse2(Launcher &L, int kind, int kid, const StageW &sw) {
  if constexpr (SCAL && MIX)
    launch(L, kind, kid, k_face2_update<true, true>, BLK, 0, sw);
  else
    with_solver<EOS_IDEAL>(L, [&](auto R) {
      launch(L, kind, kid, k_face2_update<SCAL, MIX, decltype(R)::value>, BLK, 0, sw);
    });
}

// prim: the stage writes the primitive state of shallow-water cells (within a
// step; the separate exo_add_flux_divergence call leaves it to exo_fill_ghosts)
void launch_update(Launcher &L, int stage, int prim) {
  int st0 = stage - 1;
  StageW sw{SS_DELTA[st0], SS_G1[st0], SS_G2[st0], SS_BETA[st0], st0 == 0 ? 1 : 0,
            prim, DFLOOR, SFLOOR};
  launch(L, K_UPD, KID_UPDATE, k_update<false>, BLK, 0, sw);
  launch(L, K_UPD_MIX, KID_UPDATE, k_update<true>, BLK, 0, sw);
  HIPCHK(hipGetLastError());
}

// the stage's update and column correction after launch_fluxes(L, true): the
// ideal gas's x2 faces and tendency in one kernel (k_face2_update), whose
// column step then forms the stage's weighted average of u; shallow water's
// update (the new state and its primitives)
void launch_update_implicit(Launcher &L, int stage) {
  int st0 = stage - 1;
  StageW sw{SS_DELTA[st0], SS_G1[st0], SS_G2[st0], SS_BETA[st0], st0 == 0 ? 1 : 0,
            1, DFLOOR, SFLOOR};
  launch(L, K_UPD_RS, KID_UPDATE, k_update<false, false, EOS_SW>, BLK, 0, sw);
  launch_fuse2<false, false>(L, K_FUSE2, KID_UPDATE, sw);
  launch_fuse2<true, false>(L, K_FUSE2_S, KID_UPDATE, sw);
  launch_fuse2<false, true>(L, K_FUSE2_MIX, KID_UPDATE, sw);
  launch_fuse2<true, true>(L, K_FUSE2_MIX_S, KID_UPDATE, sw);
  launch(L, K_IMPL_AVG, KID_IMPLICIT, k_implicit<true>, 64, 0, sw);
  launch(L, K_IMPL3_AVG, KID_IMPLICIT, k_implicit3<true>, 64, 0, sw);
  launch(L, K_ADD_DU_AVG, KID_IMPLICIT, k_add_du<true>, BLK, 0, sw);
  HIPCHK(hipGetLastError());
}

void launch_implicit(Launcher &L, double beta) {
  StageW sw{};
  sw.beta = beta;
  launch(L, K_IMPL, KID_IMPLICIT, k_implicit<false>, 64, 0, sw);
  launch(L, K_IMPL3, KID_IMPLICIT, k_implicit3<false>, 64, 0, sw);
  launch(L, K_ADD_DU, KID_IMPLICIT, k_add_du<false>, BLK, 0, sw);
  HIPCHK(hipGetLastError());
}

void launch_dt(Launcher &L, double dt_fixed) {
  if (dt_fixed <= 0.0) {
    launch(L, K_ONE, -1, k_dt_reset, 64, 0);
    launch(L, K_DT, KID_DT, k_dt, BLK, 0);
  }
  launch(L, K_ONE, -1, k_dt_finish, 64, 0, dt_fixed);
  HIPCHK(hipGetLastError());
}

// ghost fill within a step: the stage update has written the primitive state
// of the interior cells of shallow-water instances
void launch_fill_ghosts_stage(Launcher &L) {
  launch(L, K_GHOST_U, KID_GHOST, k_ghost, BLK, 0, 0);
  launch(L, K_GHOST_S, KID_GHOST, k_ghost, BLK, 0, 1);
  launch(L, K_PRIM_REST, KID_PRIM, k_prim<EOS_IDEAL>, BLK, 0, DFLOOR, SFLOOR);
  launch(L, K_PRIM_GHOST, KID_PRIM, k_prim_ghost, BLK, 0, DFLOOR, SFLOOR);
  HIPCHK(hipGetLastError());
}

void run_stage(Launcher &L, int stage) {
  launch_fluxes(L, true);
  launch_update_implicit(L, stage);
  launch_fill_ghosts_stage(L);
}

void run_step(Launcher &L, double dt_fixed) {
  launch_dt(L, dt_fixed);
  for (int stg = 1; stg <= 3; ++stg) run_stage(L, stg);
}

// compensated conserved totals of every instance into out[ninst][5] (device)
void launch_sums(Launcher &L, double *out) {
  launch(L, K_CELLS, -1, k_sums, BLK, 0);
  launch(L, K_ONE, -1, k_sums_final, BLK, 0, (const int *)(L.d_bstart + (size_t)K_CELLS * (L.ninst + 1)),
         out);
  HIPCHK(hipGetLastError());
}

// Autotuner of the persistent launch width: for each launch kind that only
// reads the conserved state (so repeating it changes nothing), time each
// candidate width and keep the fastest.  Widths are multiples of the CU count.
void tune_launcher(Launcher &L, int reps) {
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  bool prof = L.prof;
  L.prof = false;
  // the fused x2-face updates read the x1 and x3 face fluxes of the current
  // state, and those without scalars write only the tendency (the ones with
  // scalars rewrite s in place and keep the plain launch)
  launch_fluxes(L);
  StageW sw{1.0, 0.0, 1.0, 1.0, 1, 0, DFLOOR, SFLOOR};
  const int kinds[] = {K_PRIM, K_GHOST_U,  K_X1,          K_X1_MIX, K_EDGE,  K_EDGE_MIX,
                       K_FACE, K_FACE_MIX, K_FACE_MIX_SW, K_DT,     K_FUSE2, K_FUSE2_MIX};
  for (int kind : kinds) {
    std::vector<int> shapes = {0};  // the default first
    if (kind == K_X1 || kind == K_X1_MIX) shapes = {0, 2, 8};
    float best = 1e30f;
    int best_cap = 0, best_shape = 0;
    for (int sh : shapes) {
      set_shape(L, kind, sh);
      if (L.nblk[kind] == 0) break;
      std::vector<int> cands = {0};
      for (int m : {1, 2, 4, 8, 16})
        if ((long long)m * ncu < L.nblk[kind]) cands.push_back(m * ncu);
      for (int c : cands) {
        L.cap[kind] = c;
        for (int r = -1; r < reps; ++r) {  // r = -1 warms up
          if (r == 0) HIPCHK(hipEventRecord(e0, L.stream));
          // each candidate launch follows a primitive conversion, as within a
          // step: its inputs come from memory just written, not from caches
          // warmed by the previous repetition (the conversion's own time is
          // the same for every candidate)
          if (kind != K_PRIM) launch_prim(L, -1);
          switch (kind) {
            case K_PRIM: launch_prim(L, -1); break;
            case K_GHOST_U: launch(L, K_GHOST_U, -1, k_ghost, BLK, 0, 0); break;
            case K_X1: launch_x1(L, K_X1, -1); break;
            case K_X1_MIX: launch_x1(L, K_X1_MIX, -1); break;
            case K_EDGE: launch(L, K_EDGE, -1, k_edge_states, BLK, 0); break;
            case K_EDGE_MIX: launch(L, K_EDGE_MIX, -1, k_edge_states_mix, BLK, 0); break;
            case K_FACE: launch_face64<1>(L, K_FACE, -1); break;
            case K_FACE_MIX: launch_face_mix<1, false>(L, K_FACE_MIX, -1); break;
            case K_FACE_MIX_SW: launch_face_mix<1, true>(L, K_FACE_MIX_SW, -1); break;
            case K_DT: launch(L, K_DT, -1, k_dt, BLK, 0); break;
            case K_FUSE2: launch_fuse2<false, false>(L, K_FUSE2, -1, sw); break;
            case K_FUSE2_MIX: launch_fuse2<false, true>(L, K_FUSE2_MIX, -1, sw); break;
          }
        }
        HIPCHK(hipEventRecord(e1, L.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best * 0.97f) {  // prefer the default launch unless clearly slower
          best = ms;
          best_cap = c;
          best_shape = sh;
        }
      }
    }
    set_shape(L, kind, best_shape);
    L.cap[kind] = best_cap;
  }
  // kinds sharing a shape with a tuned one (the other update-shaped kinds keep
  // the plain launch: they rewrite the state and cannot be repeated)
  L.cap[K_GHOST_S] = L.cap[K_GHOST_U];
  L.cap[K_FACE_NF] = L.cap[K_FACE];
  L.prof = prof;
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipGetLastError());
}

// copy between the host MeshBlock layout (p, n, k, j, i) and the device
// layout (n, p, k, i, j)
void host_to_dev(exo_handle *h, const double *src, double *dst, int nvar) {
  Grid &g = h->g;
  std::vector<double> buf((size_t)nvar * g.vstride, 0.);
  for (int p = 0; p < 6; ++p)
    for (int n = 0; n < nvar; ++n)
      for (int k = 0; k < g.nc3; ++k)
        for (int j = 0; j < g.nc2; ++j)
          for (int i = 0; i < g.nc1; ++i)
            buf[g.c(n, p, k, i, j)] =
                src[((((size_t)p * nvar + n) * g.nc3 + k) * g.nc2 + j) * g.nc1 + i];
  HIPCHK(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
}

void dev_to_host(exo_handle *h, const double *src, double *dst, int nvar) {
  Grid &g = h->g;
  std::vector<double> buf((size_t)nvar * g.vstride);
  HIPCHK(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int p = 0; p < 6; ++p)
    for (int n = 0; n < nvar; ++n)
      for (int k = 0; k < g.nc3; ++k)
        for (int j = 0; j < g.nc2; ++j)
          for (int i = 0; i < g.nc1; ++i)
            dst[((((size_t)p * nvar + n) * g.nc3 + k) * g.nc2 + j) * g.nc1 + i] =
                buf[g.c(n, p, k, i, j)];
}

// Per-level reference of the mixed path: the horizontal mean of the uploaded
// state at each level (pressure and density for the ideal gas, geopotential
// for shallow water).  Any level-constant reference gives the same fluxes up to
// FP32 rounding; a close one keeps the perturbations small.
void set_reference(exo_handle *h, const double *u) {
  Grid &g = h->g;
  int nv = g.nv, nc1 = g.nc1, nc2 = g.nc2, nc3 = g.nc3;
  std::vector<double> ref(3 * (size_t)nc1, 0.), pr(nc1, 0.), rh(nc1, 0.);
  double gm1 = g.gamma - 1.0;
  for (int i = g.is; i <= g.ie; ++i) {
    double sp = 0., sr = 0.;
    for (int p = 0; p < 6; ++p)
      for (int k = g.ks; k <= g.ke; ++k)
        for (int j = g.js; j <= g.je; ++j) {
          auto U = [&](int n) { return u[((((size_t)p * nv + n) * nc3 + k) * nc2 + j) * nc1 + i]; };
          double d = U(IDN);
          sr += d;
          if (g.eos == EOS_IDEAL)
            sp += gm1 * (U(IEN) - 0.5 * (U(IVX) * U(IVX) + U(IVY) * U(IVY) + U(IVZ) * U(IVZ)) / d);
        }
    double cnt = 6.0 * g.N * g.N;
    rh[i] = sr / cnt;
    pr[i] = sp / cnt;
  }
  for (int i = 0; i < nc1; ++i) {
    int ii = std::min(std::max(i, g.is), g.ie);
    if (g.eos == EOS_SW) {
      ref[i] = 0.5 * rh[ii] * rh[ii];  // split "pressure" h^2/2
      ref[nc1 + i] = rh[ii];           // geopotential reference
    } else {
      ref[i] = pr[ii];
      ref[nc1 + i] = pr[ii];
      ref[2 * nc1 + i] = rh[ii];
    }
  }
  HIPCHK(hipMemcpyAsync(h->st.lref, ref.data(), ref.size() * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
}

void prof_begin(Launcher &L) {
  HIPCHK(hipStreamSynchronize(L.stream));
  L.ev_used = 0;
  L.ev_kind.clear();
  L.prof = true;
}

void prof_end(Launcher &L, double *ms, int *counts, int nkinds) {
  HIPCHK(hipStreamSynchronize(L.stream));
  L.prof = false;
  for (int q = 0; q < nkinds; ++q) {
    if (ms) ms[q] = 0.;
    if (counts) counts[q] = 0;
  }
  for (size_t e = 0; e < L.ev_kind.size(); ++e) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, L.ev_pool[2 * e], L.ev_pool[2 * e + 1]));
    int kid = L.ev_kind[e];
    if (kid < nkinds) {
      if (ms) ms[kid] += t;
      if (counts) counts[kid] += 1;
    }
  }
  L.ev_used = 0;
  L.ev_kind.clear();
}

template <typename T, typename F>
int guarded(T *h, F &&fn) {
  if (!h) return EXO_ERR_ARG;
  try {
    fn();
    return EXO_OK;
  } catch (const std::string &m) {
    h->err = m;
    return m.rfind("HIP", 0) == 0 ? EXO_ERR_HIP : EXO_ERR_ARG;
  } catch (const std::exception &e) {
    h->err = e.what();
    return EXO_ERR_NOMEM;
  }
}

// Coefficient scratch of exo_implicit_columns_noalloc: a fixed array in
// device memory (each GPU the library runs on holds its own copy until the
// process ends), so that entry point allocates nothing.  Every call on one GPU
// shares it: calls that overlap on one GPU race.  Nothing in it needs clearing:
// each column's forward sweep writes every level it holds before the backward
// substitution reads it.
static __device__ __align__(256) double scratch[EXO_NOALLOC_SCRATCH_DOUBLES];

// the device address of this GPU's copy of scratch (the array's name in host
// code is not a device address); looked up once per device, with that device
// current
double *scratch_ptr() {
  static std::mutex mu;
  static double *ptr[64] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::string("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = (double *)p;
  }
  return ptr[dev];
}

// Output sets: fields per cell, the device buffer they go to, the launch.
size_t output_floats(const Grid &g, int set, int ghost) {
  int G = ghost ? NG : 0;
  size_t n1 = g.nx1 + 2 * G, n2 = g.N + 2 * G;
  return (size_t)6 * out_names(set, g.eos, g.nt).size() * n1 * n2 * n2;
}

// writes the fields of one output set into h->d_out (stream-ordered after
// every launch before it); returns the number of fields
int output_fields_dev(exo_handle *h, int set, int ghost) {
  Grid &g = h->g;
  int nq = (int)out_names(set, g.eos, g.nt).size();
  if (nq > OUT_MAX) throw std::string("too many output fields (at most 24)");
  size_t need = output_floats(g, set, ghost);
  if (need > h->out_cap) {
    if (h->d_out) HIPCHK(hipFree(h->d_out));
    h->d_out = nullptr;
    h->out_cap = 0;
    HIPCHK(hipMalloc(&h->d_out, need * sizeof(float)));
    h->out_cap = need;
  }
  OutMap m{};
  size_t tp = (size_t)6 * g.nc3 * g.nc2;
  if (set == OUT_UOV) {
    if (!h->d_uov) {
      std::vector<double> t = uov_table(h->setup);
      h->d_uov = (double *)h->dalloc(t.size() * sizeof(double));
      HIPCHK(hipMemcpyAsync(h->d_uov, t.data(), t.size() * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
    }
    const int ops[5] = {OP_TAB, OP_TAB, OP_U, OP_V, OP_TAB};
    const double *src[5] = {h->d_uov, h->d_uov + tp, nullptr, nullptr, h->d_uov + 2 * tp};
    for (int q = 0; q < 5; ++q) {
      m.op[q] = ops[q];
      m.src[q] = src[q];
    }
  } else {
    const double *base = set == OUT_PRIM ? h->st.w : h->st.u;
    const double *sbase = set == OUT_PRIM ? h->st.r : h->st.s;
    int q = 0;
    m.src[q++] = base + IDN * g.vstride;
    if (g.eos == EOS_IDEAL) m.src[q++] = base + IPR * g.vstride;
    for (int d = IVX; d <= IVZ; ++d) {
      m.flip[q] = d == IVX;
      m.src[q++] = base + d * g.vstride;
    }
    for (int t = 0; t < g.nt; ++t) m.src[q++] = sbase + t * g.vstride;
    for (int e = 0; e < q; ++e) m.op[e] = OP_COPY;
  }
  int G = ghost ? NG : 0;
  unsigned n2 = g.N + 2 * G, nt = (n2 + 63) / 64;
  dim3 grid(nt * nt, g.nx1 + 2 * G, 6 * nq);
  Launcher &L = h->L;
  if (L.prof) {
    L.ev_kind.push_back(KID_OUTPUT);
    HIPCHK(hipEventRecord(next_event(L), L.stream));
  }
  k_out_fields<<<grid, dim3(256), 0, h->stream>>>(g, m, h->st.w + IVY * g.vstride,
                                                   h->st.w + IVZ * g.vstride, h->d_uov, nq,
                                                   ghost, h->d_out);
  HIPCHK(hipGetLastError());
  if (L.prof) HIPCHK(hipEventRecord(next_event(L), L.stream));
  return nq;
}

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int exo_create(const exo_config *cfg, exo_handle **out) {
  if (!cfg || !out) return EXO_ERR_ARG;
  *out = nullptr;
  exo_handle *h = new exo_handle();
  std::string m = build_setup(*cfg, &h->setup);
  if (!m.empty()) {
    fprintf(stderr, "exo_create: %s\n", m.c_str());
    delete h;
    return EXO_ERR_ARG;
  }
  if (cfg->precision != EXO_PREC_FP64 && cfg->precision != EXO_PREC_MIXED) {
    delete h;
    return EXO_ERR_ARG;
  }
  try {
    init_handle(h);
  } catch (const std::string &e) {
    fprintf(stderr, "exo_create: %s\n", e.c_str());
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return EXO_ERR_HIP;
  }
  *out = h;
  return EXO_OK;
}

int exo_destroy(exo_handle *h) {
  if (!h) return EXO_ERR_ARG;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (hipEvent_t e : h->L.ev_pool) (void)hipEventDestroy(e);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->d_out) (void)hipFree(h->d_out);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return EXO_OK;
}

const char *exo_last_error(const exo_handle *h) { return h ? h->err.c_str() : "null handle"; }
int exo_nvar(const exo_handle *h) { return h ? h->g.nv : 0; }
size_t exo_state_doubles(const exo_handle *h) {
  return h ? (size_t)6 * h->g.nv * h->g.nc3 * h->g.nc2 * h->g.nc1 : 0;
}

int exo_upload_state(exo_handle *h, const double *u, const double *s) {
  return guarded(h, [&] {
    if (!u) throw std::string("null state");
    host_to_dev(h, u, h->st.u, h->g.nv);
    if (h->g.nt) {
      if (!s) throw std::string("scalars required (ntracer > 0)");
      host_to_dev(h, s, h->st.s, h->g.nt);
    }
    double dt0[2] = {0., DBL_MAX * 0.4};
    HIPCHK(hipMemcpyAsync(h->st.dtv, dt0, sizeof(dt0), hipMemcpyHostToDevice, h->stream));
    if (h->mixed) set_reference(h, u);
    HIPCHK(hipMemsetAsync(h->st.bad, 0, sizeof(int), h->stream));
    launch_fill_ghosts(h->L);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_state = true;
  });
}

int exo_download_state(exo_handle *h, double *u, double *s) {
  return guarded(h, [&] {
    if (!u) throw std::string("null output");
    dev_to_host(h, h->st.u, u, h->g.nv);
    if (h->g.nt && s) dev_to_host(h, h->st.s, s, h->g.nt);
  });
}

int exo_download_prim(exo_handle *h, double *w) {
  return guarded(h, [&] {
    if (!w) throw std::string("null output");
    dev_to_host(h, h->st.w, w, h->g.nv);
  });
}

int exo_calc_fluxes(exo_handle *h) {
  return guarded(h, [&] {
    if (!h->have_state) throw std::string("no state uploaded");
    launch_fluxes(h->L);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_download_flux(exo_handle *h, int dir, double *out) {
  return guarded(h, [&] {
    if (dir < 0 || dir > 2 || !out) throw std::string("bad flux direction");
    if (h->mixed)
      throw std::string("the mixed form's face fluxes are not the reference's FP64 fluxes "
                        "(FP32 perturbation or delta parts); download fluxes from an "
                        "EXO_PREC_FP64 handle");
    Grid &g = h->g;
    std::vector<double> buf((size_t)g.nv * g.fvstride);
    HIPCHK(hipMemcpyAsync(buf.data(), h->st.flx[dir], buf.size() * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    int n1 = g.nc1 + (dir == 0), n2 = g.nc2 + (dir == 1), n3 = g.nc3 + (dir == 2);
    for (int p = 0; p < 6; ++p)
      for (int n = 0; n < g.nv; ++n)
        for (int k = 0; k < n3; ++k)
          for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i)
              out[((((size_t)p * g.nv + n) * n3 + k) * n2 + j) * n1 + i] = buf[g.f(n, p, k, i, j)];
  });
}

int exo_add_flux_divergence(exo_handle *h, double wght) {
  return guarded(h, [&] {
    // stage-1 weights with beta * dt = wght: u1 <- u, u <- u1, then the update
    double dtv = wght;
    HIPCHK(hipMemcpyAsync(h->st.dtv, &dtv, sizeof(double), hipMemcpyHostToDevice, h->stream));
    launch_update(h->L, 1, 0);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_integrate_hydro(exo_handle *h, int stage, double dt) {
  return guarded(h, [&] {
    if (stage < 1 || stage > 3) throw std::string("stage must be 1..3");
    // the stage's weighted average of u and u1, then the update with
    // wght = beta(stage) dt (the kernel forms beta dt from the dt it reads)
    double dtv = dt;
    HIPCHK(hipMemcpyAsync(h->st.dtv, &dtv, sizeof(double), hipMemcpyHostToDevice, h->stream));
    launch_update(h->L, stage, 0);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_fill_ghosts(exo_handle *h) {
  return guarded(h, [&] {
    launch_fill_ghosts(h->L);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_implicit_correct(exo_handle *h, double dt_stage) {
  return guarded(h, [&] {
    double dtv = dt_stage;
    HIPCHK(hipMemcpyAsync(h->st.dtv, &dtv, sizeof(double), hipMemcpyHostToDevice, h->stream));
    launch_implicit(h->L, 1.0);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_new_dt(exo_handle *h, double *dt) {
  return guarded(h, [&] {
    launch_dt(h->L, 0.0);
    double v[2];
    HIPCHK(hipMemcpyAsync(v, h->st.dtv, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (dt) *dt = v[0];
  });
}

int exo_stage(exo_handle *h, int stage, double dt) {
  return guarded(h, [&] {
    if (stage < 1 || stage > 3) throw std::string("stage must be 1..3");
    double dtv = dt;
    HIPCHK(hipMemcpyAsync(h->st.dtv, &dtv, sizeof(double), hipMemcpyHostToDevice, h->stream));
    run_stage(h->L, stage);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

}  // extern "C"

namespace {
// the step graph of a handle (captured on first use, rebuilt when the dt mode
// or the launch widths change); kernel nodes only
void handle_graph(exo_handle *h, double dt_fixed) {
  if (h->exec && h->graph_dt_fixed == dt_fixed) return;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->exec = nullptr;
  h->graph = nullptr;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  try {
    run_step(h->L, dt_fixed);
  } catch (...) {
    hipGraph_t gr = nullptr;
    (void)hipStreamEndCapture(h->stream, &gr);
    if (gr) (void)hipGraphDestroy(gr);
    throw;
  }
  HIPCHK(hipStreamEndCapture(h->stream, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  h->graph_dt_fixed = dt_fixed;
}

// Mesh::SaveAllStates / LoadAllStates (src/mesh_redo.cpp:12-24 over
// MeshBlock::Impl::SaveAllStates, impl.cpp:306-332): the conserved and
// primitive state, here with the passive scalars and the time-step state,
// copied within HBM
size_t snap_doubles(const exo_handle *h) {
  size_t fw = h->st.fw ? ((size_t)(h->g.nv + h->g.nt) * h->g.vstride + 1) / 2 : 0;
  return (size_t)(2 * h->g.nv + 2 * h->g.nt) * h->g.vstride + 2 + fw;
}

void snap_copy(exo_handle *h, bool save) {
  const Grid &g = h->g;
  double *d = h->d_snap;
  auto cp = [&](double *state, size_t n) {
    HIPCHK(hipMemcpyAsync(save ? d : state, save ? state : d, n * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
    d += n;
  };
  cp(h->st.u, (size_t)g.nv * g.vstride);
  cp(h->st.w, (size_t)g.nv * g.vstride);
  if (g.nt) {
    cp(h->st.s, (size_t)g.nt * g.vstride);
    cp(h->st.r, (size_t)g.nt * g.vstride);
  }
  cp(h->st.dtv, 2);
  if (h->st.fw) {  // the FP32 perturbations the face kernels read
    size_t bytes = (size_t)(g.nv + g.nt) * g.vstride * sizeof(float);
    HIPCHK(hipMemcpyAsync(save ? (void *)d : (void *)h->st.fw, save ? (void *)h->st.fw : (void *)d,
                          bytes, hipMemcpyDeviceToDevice, h->stream));
  }
}

void save_states(exo_handle *h) {
  if (!h->d_snap) h->d_snap = (double *)h->dalloc(snap_doubles(h) * sizeof(double));
  snap_copy(h, true);
}

void load_states(exo_handle *h) {
  if (!h->d_snap) throw std::string("no saved state");
  snap_copy(h, false);
}

bool state_valid(exo_handle *h) {
  int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, h->st.bad, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return bad == 0;
}

// ---------------------------------------------------------------------------
// checkpoint files: everything a restart needs to continue bit for bit
constexpr char CKPT_MAGIC[8] = {'E', 'X', 'O', 'C', 'K', 'P', 'T', '1'};

struct CkptMeta {
  double time, dt, dt_prev;
  int64_t ncycle;
};

void put_cfg(std::vector<double> &m, const exo_config &c) {
  for (int v : {c.nx, c.nx1, c.eos, c.rsolver, c.implicit_flag, c.ntracer, c.coriolis,
                c.cfl_x1, c.precision})
    m.push_back((double)v);
  for (double v : {c.x1min, c.x1max, c.gamma, c.grav, c.omega, c.cfl}) m.push_back(v);
}

// FNV-1a over the payload's 64-bit words
uint64_t fnv64(const double *a, size_t n, uint64_t hsh) {
  for (size_t e = 0; e < n; ++e) {
    uint64_t x;
    std::memcpy(&x, a + e, 8);
    hsh = (hsh ^ x) * 0x100000001b3ull;
  }
  return hsh;
}
}  // namespace
extern "C" {

int exo_step(exo_handle *h, int nsteps, double dt_fixed, double *dt_last) {
  return guarded(h, [&] {
    if (!h->have_state) throw std::string("no state uploaded");
    if (nsteps > 1 && !h->L.prof) {  // replay a captured step (no per-kernel timing)
      handle_graph(h, dt_fixed);
      for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(h->exec, h->stream));
    } else {
      for (int s = 0; s < nsteps; ++s) run_step(h->L, dt_fixed);
    }
    if (dt_last) {
      double v[2];
      HIPCHK(hipMemcpyAsync(v, h->st.dtv, sizeof(v), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      *dt_last = v[0];
    }
  });
}

// ---------------------------------------------------------------------------
// redo / rollback (main.cpp.in:126-160, src/mesh_redo.cpp:12-45)
int exo_state_valid(exo_handle *h, int *valid) {
  return guarded(h, [&] {
    if (!valid) throw std::string("null output");
    *valid = state_valid(h) ? 1 : 0;
  });
}

int exo_save_states(exo_handle *h) {
  return guarded(h, [&] {
    if (!h->have_state) throw std::string("no state uploaded");
    save_states(h);
  });
}

int exo_load_states(exo_handle *h) {
  return guarded(h, [&] {
    load_states(h);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_step_redo(exo_handle *h, int nsteps, int max_redo, double *dt_last, int *redos) {
  return guarded(h, [&] {
    if (!h->have_state) throw std::string("no state uploaded");
    bool redo = false;
    double dt_redo = 0.;
    int done = 0, n_redo = 0;
    if (redos) *redos = 0;
    while (done < nsteps) {
      if (!redo && max_redo > 0) save_states(h);
      if (redo) {
        // Mesh::DecreaseTimeStep: the halved dt is the step's dt, and the
        // next step may grow to twice it (Mesh::NewTimeStep)
        double v[2] = {dt_redo, dt_redo};
        HIPCHK(hipMemcpyAsync(h->st.dtv, v, sizeof(v), hipMemcpyHostToDevice, h->stream));
        run_step(h->L, dt_redo);
      } else {
        run_step(h->L, 0.0);
      }
      if (max_redo > 0) {
        if (state_valid(h)) {
          redo = false;
        } else {
          double v[2];
          HIPCHK(hipMemcpyAsync(v, h->st.dtv, sizeof(v), hipMemcpyDeviceToHost, h->stream));
          HIPCHK(hipStreamSynchronize(h->stream));
          load_states(h);
          dt_redo = v[0] / 2;
          redo = true;
          ++n_redo;
          if (redos) *redos = n_redo;
          if (h->redo_attempts++ > max_redo)
            throw std::string("Too many attempts to redo the step");
          continue;
        }
      }
      ++done;
    }
    if (dt_last) {
      double v[2];
      HIPCHK(hipMemcpyAsync(v, h->st.dtv, sizeof(v), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      *dt_last = v[0];
    }
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_get_dt_state(exo_handle *h, double *dt, double *dt_prev) {
  return guarded(h, [&] {
    double v[2];
    HIPCHK(hipMemcpyAsync(v, h->st.dtv, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (dt) *dt = v[0];
    if (dt_prev) *dt_prev = v[1];
  });
}

int exo_set_dt_state(exo_handle *h, double dt, double dt_prev) {
  return guarded(h, [&] {
    if (!(dt_prev > 0.)) throw std::string("dt_prev must be positive");
    double v[2] = {dt, dt_prev};
    HIPCHK(hipMemcpyAsync(h->st.dtv, v, sizeof(v), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

// ---------------------------------------------------------------------------
// checkpoint / restart
int exo_save_checkpoint(exo_handle *h, const char *path, double time, int64_t ncycle) {
  return guarded(h, [&] {
    if (!path) throw std::string("null path");
    if (!h->have_state) throw std::string("no state uploaded");
    const Grid &g = h->g;
    size_t nu = (size_t)6 * g.nv * g.nc3 * g.nc2 * g.nc1;
    size_t ns = (size_t)6 * g.nt * g.nc3 * g.nc2 * g.nc1;
    size_t nl = h->mixed ? 3 * (size_t)g.nc1 : 0;
    std::vector<double> pay(nu + ns + nl);
    dev_to_host(h, h->st.u, pay.data(), g.nv);
    if (g.nt) dev_to_host(h, h->st.s, pay.data() + nu, g.nt);
    if (nl)
      HIPCHK(hipMemcpyAsync(pay.data() + nu + ns, h->st.lref, nl * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    double dtv[2];
    HIPCHK(hipMemcpyAsync(dtv, h->st.dtv, sizeof(dtv), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> meta;
    put_cfg(meta, h->setup.cfg);
    meta.push_back(time);
    meta.push_back(dtv[0]);
    meta.push_back(dtv[1]);
    double nc;
    std::memcpy(&nc, &ncycle, 8);  // the cycle count's bits
    meta.push_back(nc);
    uint64_t counts[3] = {nu, ns, nl};
    uint64_t sum = fnv64(pay.data(), pay.size(), fnv64(meta.data(), meta.size(),
                                                        0xcbf29ce484222325ull));
    std::string tmp = std::string(path) + ".tmp";
    std::FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) throw std::string("cannot open ") + tmp + " for writing";
    uint64_t nmeta = meta.size();
    bool ok = std::fwrite(CKPT_MAGIC, 1, 8, fp) == 8 &&
              std::fwrite(&nmeta, 8, 1, fp) == 1 &&
              std::fwrite(meta.data(), 8, nmeta, fp) == nmeta &&
              std::fwrite(counts, 8, 3, fp) == 3 &&
              std::fwrite(pay.data(), 8, pay.size(), fp) == pay.size() &&
              std::fwrite(&sum, 8, 1, fp) == 1;
    if (std::fclose(fp) != 0) ok = false;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
      std::remove(tmp.c_str());
      throw std::string("write error on ") + path;
    }
  });
}

int exo_load_checkpoint(exo_handle *h, const char *path, double *time, int64_t *ncycle) {
  return guarded(h, [&] {
    if (!path) throw std::string("null path");
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) throw std::string("cannot open ") + path;
    std::vector<double> meta(19), pay;
    uint64_t nmeta = 0, counts[3] = {0, 0, 0}, sum = 0;
    char magic[8];
    bool ok = std::fread(magic, 1, 8, fp) == 8 && !std::memcmp(magic, CKPT_MAGIC, 8) &&
              std::fread(&nmeta, 8, 1, fp) == 1 && nmeta == meta.size() &&
              std::fread(meta.data(), 8, nmeta, fp) == nmeta && std::fread(counts, 8, 3, fp) == 3;
    if (!ok) {
      std::fclose(fp);
      throw std::string("not a checkpoint file: ") + path;
    }
    const Grid &g = h->g;
    size_t nu = (size_t)6 * g.nv * g.nc3 * g.nc2 * g.nc1;
    size_t ns = (size_t)6 * g.nt * g.nc3 * g.nc2 * g.nc1;
    std::vector<double> mine;
    put_cfg(mine, h->setup.cfg);
    // the configuration must match except for the precision
    for (int q = 0; q < 15; ++q)
      if (q != 8 && std::memcmp(&mine[q], &meta[q], 8) != 0) {
        std::fclose(fp);
        throw std::string("checkpoint of another configuration: ") + path;
      }
    ok = counts[0] == nu && counts[1] == ns && (counts[2] == 0 || counts[2] == 3 * (size_t)g.nc1);
    if (ok) {
      pay.resize(counts[0] + counts[1] + counts[2]);
      ok = std::fread(pay.data(), 8, pay.size(), fp) == pay.size() &&
           std::fread(&sum, 8, 1, fp) == 1;
    }
    std::fclose(fp);
    if (!ok) throw std::string("not a complete checkpoint file: ") + path;
    if (fnv64(pay.data(), pay.size(), fnv64(meta.data(), meta.size(), 0xcbf29ce484222325ull)) !=
        sum)
      throw std::string("checkpoint checksum mismatch: ") + path;
    host_to_dev(h, pay.data(), h->st.u, g.nv);
    if (g.nt) host_to_dev(h, pay.data() + nu, h->st.s, g.nt);
    double dtv[2] = {meta[16], meta[17]};
    HIPCHK(hipMemcpyAsync(h->st.dtv, dtv, sizeof(dtv), hipMemcpyHostToDevice, h->stream));
    if (h->mixed) {
      if (counts[2])  // the reference levels of the run that wrote it
        HIPCHK(hipMemcpyAsync(h->st.lref, pay.data() + nu + ns, counts[2] * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
      else
        set_reference(h, pay.data());
    }
    HIPCHK(hipMemsetAsync(h->st.bad, 0, sizeof(int), h->stream));
    launch_fill_ghosts(h->L);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_state = true;
    if (time) *time = meta[15];
    if (ncycle) std::memcpy(ncycle, &meta[18], 8);
  });
}

int exo_profile_begin(exo_handle *h) {
  return guarded(h, [&] { prof_begin(h->L); });
}

int exo_profile_end(exo_handle *h, double *ms, int *counts, int nkinds) {
  return guarded(h, [&] { prof_end(h->L, ms, counts, nkinds); });
}

int exo_sync(exo_handle *h) {
  return guarded(h, [&] { HIPCHK(hipStreamSynchronize(h->stream)); });
}

int exo_init_state(exo_handle *h, int kind, uint64_t seed, const double *params, int nparams,
                   double *u_out, double *s_out) {
  return guarded(h, [&] {
    if (!u_out) throw std::string("null output");
    std::string m = init_state(h->setup, kind, seed, params, nparams, u_out, s_out);
    if (!m.empty()) throw m;
  });
}

int exo_riemann_face(exo_handle *h, int dir, int p, int k, int j, const double *wl,
                     const double *wr, double *flx) {
  return guarded(h, [&] {
    if (dir < 0 || dir > 2 || p < 0 || p > 5 || k < 0 || k >= h->g.nc3 || j < 0 ||
        j >= h->g.nc2)
      throw std::string("bad face");
    double *d;
    HIPCHK(hipMalloc(&d, 15 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(d, wl, h->g.nv * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(d + 5, wr, h->g.nv * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_riemann_face, dim3(1), dim3(1), 0, h->stream, h->g, dir, p, k, j, d,
                       d + 5, d + 10);
    HIPCHK(hipMemcpyAsync(flx, d + 10, h->g.nv * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(d));
  });
}

int exo_implicit_column(exo_handle *h, int nx1, const double *w_col, double *du_col,
                        const double *area1, const double *vol, double dt) {
  return guarded(h, [&] {
    if (h->g.eos != EOS_IDEAL) throw std::string("the column solve needs the ideal-gas EOS");
    if (nx1 < 1 || nx1 + 2 * NG > IMPL_MAXNC1) throw std::string("bad column length");
    if (!w_col || !du_col || !area1 || !vol) throw std::string("null column array");
    int nc1 = nx1 + 2 * NG;
    size_t nw = (size_t)5 * nc1, ncoef = (size_t)nx1 * 15 * 64 * 2;
    double *d;
    HIPCHK(hipMalloc(&d, (ncoef + 2 * nw + 2 * (size_t)nx1 + 1) * sizeof(double)));
    double *dw = d + ncoef, *ddu = dw + nw, *da = ddu + nw, *dv = da + nx1 + 1;
    HIPCHK(hipMemcpyAsync(dw, w_col, nw * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(ddu, du_col, nw * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(da, area1, (nx1 + 1) * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    HIPCHK(hipMemcpyAsync(dv, vol, nx1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ColIOShim io{dw, ddu, da, dv, nc1, NG};
    hipLaunchKernelGGL(k_implicit_columns, dim3(1), dim3(64), 0, h->stream, io, 1, nx1, dt,
                       h->g.gamma, h->g.grav, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(du_col, ddu, nw * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(d));
  });
}

int exo_implicit_columns_noalloc(exo_handle *h, int ncol, int nx1, const double *w_cols,
                                 double *du_cols, const double *area1, const double *vol,
                                 double dt) {
  return guarded(h, [&] {
    if (h->g.eos != EOS_IDEAL) throw std::string("the column solve needs the ideal-gas EOS");
    if (ncol < 1 || nx1 < 1 || nx1 + 2 * NG > IMPL_MAXNC1)
      throw std::string("bad column count or length");
    if (!w_cols || !du_cols || !area1 || !vol) throw std::string("null column array");
    size_t nblk = ((size_t)ncol + 63) / 64;
    if (nblk * nx1 * 15 * 64 * 2 > (size_t)EXO_NOALLOC_SCRATCH_DOUBLES)
      throw std::string("columns need ") + std::to_string(nblk * nx1 * 15 * 64 * 2) +
          " doubles of coefficient scratch, above the fixed " +
          std::to_string(EXO_NOALLOC_SCRATCH_DOUBLES);
    double *scr = scratch_ptr();
    ColIOShim io{w_cols, du_cols, area1, vol, nx1 + 2 * NG, NG};
    hipLaunchKernelGGL(k_implicit_columns, dim3((unsigned)nblk), dim3(64), 0, h->stream, io,
                       ncol, nx1, dt, h->g.gamma, h->g.grav, scr);
    HIPCHK(hipGetLastError());
  });
}

int exo_weno5_batch(exo_handle *h, const double *in5, double *out, int n) {
  return guarded(h, [&] {
    if (n <= 0) return;
    double *d;
    HIPCHK(hipMalloc(&d, (size_t)6 * n * sizeof(double)));
    HIPCHK(hipMemcpyAsync(d, in5, (size_t)5 * n * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    hipLaunchKernelGGL(k_weno_batch, dim3(grid1(n)), dim3(BLK), 0, h->stream, d, d + 5 * (size_t)n,
                       n);
    HIPCHK(hipMemcpyAsync(out, d + 5 * (size_t)n, (size_t)n * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(d));
  });
}

int exo_math_batch(exo_handle *h, int op, const double *x, double y, double *out, int n) {
  return guarded(h, [&] {
    if (n <= 0) return;
    if (op < 0 || op > 4) throw std::string("bad op");
    size_t nin = (size_t)n * (op >= 3 ? 2 : 1);
    double *d;
    HIPCHK(hipMalloc(&d, (nin + n) * sizeof(double)));
    HIPCHK(hipMemcpyAsync(d, x, nin * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_math_batch, dim3(grid1(n)), dim3(BLK), 0, h->stream, op, d, y, d + nin,
                       n);
    HIPCHK(hipMemcpyAsync(out, d + nin, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(d));
  });
}

int exo_init_state_cfg(const exo_config *cfg, int kind, uint64_t seed, const double *params,
                       int nparams, double *u_out, double *s_out) {
  if (!cfg || !u_out) return EXO_ERR_ARG;
  Setup s;
  if (!build_setup(*cfg, &s).empty()) return EXO_ERR_ARG;
  std::string m = init_state(s, kind, seed, params, nparams, u_out, s_out);
  if (!m.empty()) {
    fprintf(stderr, "exo_init_state_cfg: %s\n", m.c_str());
    return EXO_ERR_ARG;
  }
  return EXO_OK;
}

// ---------------------------------------------------------------------------
// output (NetcdfOutput::WriteOutputFile, src/outputs/netcdf.cpp:34-469)
int exo_output_names(const exo_config *cfg, const char *variable, char *names, size_t cap,
                     int *nvar) {
  int set = out_set_from_name(variable);
  if (!cfg || !nvar || set < 0) return EXO_ERR_ARG;
  std::vector<std::string> n = out_names(set, cfg->eos, cfg->ntracer);
  std::string joined;
  for (auto &x : n) joined += (joined.empty() ? "" : ",") + x;
  *nvar = (int)n.size();
  if (names) {
    if (cap < joined.size() + 1) return EXO_ERR_ARG;
    std::memcpy(names, joined.c_str(), joined.size() + 1);
  }
  return EXO_OK;
}

int exo_output_fields(exo_handle *h, const char *variable, int ghost_zones, float *out,
                      size_t cap_floats) {
  return guarded(h, [&] {
    int set = out_set_from_name(variable);
    if (set < 0) throw std::string("unknown output variable (prim, cons or uov)");
    if (!h->have_state) throw std::string("no state uploaded");
    if (!out) throw std::string("null output");
    size_t need = output_floats(h->g, set, ghost_zones != 0);
    if (cap_floats < need)
      throw std::string("output buffer too small: ") + std::to_string(need) + " floats needed";
    output_fields_dev(h, set, ghost_zones != 0);
    HIPCHK(hipMemcpyAsync(out, h->d_out, need * sizeof(float), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int exo_write_netcdf(exo_handle *h, const char *basename, const char *variable,
                     const char *file_id, int file_number, double time, int blocks_per_side,
                     int ghost_zones) {
  return guarded(h, [&] {
    int set = out_set_from_name(variable);
    if (set < 0) throw std::string("unknown output variable (prim, cons or uov)");
    if (!basename) throw std::string("null basename");
    if (!h->have_state) throw std::string("no state uploaded");
    size_t need = output_floats(h->g, set, ghost_zones != 0);
    std::vector<float> buf(need);
    output_fields_dev(h, set, ghost_zones != 0);
    HIPCHK(hipMemcpyAsync(buf.data(), h->d_out, need * sizeof(float), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::string m = write_output_files(h->setup, basename, set, file_id ? file_id : variable,
                                       file_number, time, blocks_per_side, ghost_zones != 0,
                                       buf.data());
    if (!m.empty()) throw m;
  });
}

int exo_write_netcdf_fields(const exo_config *cfg, const char *basename, const char *variable,
                            const char *file_id, int file_number, double time,
                            int blocks_per_side, int ghost_zones, const float *fields) {
  int set = out_set_from_name(variable);
  if (!cfg || !basename || !fields || set < 0) return EXO_ERR_ARG;
  Setup s;
  if (!build_setup(*cfg, &s).empty()) return EXO_ERR_ARG;
  std::string m = write_output_files(s, basename, set, file_id ? file_id : variable, file_number,
                                     time, blocks_per_side, ghost_zones != 0, fields);
  if (m.empty()) return EXO_OK;
  return m.rfind("cannot open", 0) == 0 || m.rfind("write error", 0) == 0 ? EXO_ERR_IO
                                                                         : EXO_ERR_ARG;
}

// Host-only introspection (no device needed): the set-up tables a handle
// would upload.  which: 0 ghost descriptors, 1 edge descriptors, 2 metric
// tables [T_NTAB][6][nc3][nc2], 3 x1f|x1v|dx1f, 4 x2f of the six panels,
// 5 x3f of the six panels, 6 per-row tangent tables [T1_NTAB][6][nc2 + 1].
int exo_describe(const exo_config *cfg, int which, void *buf, size_t cap, size_t *count) {
  if (!cfg || !count) return EXO_ERR_ARG;
  Setup s;
  std::string m = build_setup(*cfg, &s);
  if (!m.empty()) return EXO_ERR_ARG;
  const void *src = nullptr;
  size_t n = 0, esz = 0;
  std::vector<double> tmp;
  switch (which) {
    case 0: src = s.ghosts.data(); n = s.ghosts.size(); esz = sizeof(GhostDesc); break;
    case 1: src = s.edges.data(); n = s.edges.size(); esz = sizeof(EdgeDesc); break;
    case 2: src = s.tab2.data(); n = s.tab2.size(); esz = sizeof(double); break;
    case 3:
      tmp = s.x1f;
      tmp.insert(tmp.end(), s.x1v.begin(), s.x1v.end());
      tmp.insert(tmp.end(), s.dx1f.begin(), s.dx1f.end());
      src = tmp.data(); n = tmp.size(); esz = sizeof(double); break;
    case 4:
    case 5:
      for (int p = 0; p < 6; ++p) {
        auto &v = which == 4 ? s.x2f[p] : s.x3f[p];
        tmp.insert(tmp.end(), v.begin(), v.end());
      }
      src = tmp.data(); n = tmp.size(); esz = sizeof(double); break;
    case 6: src = s.tab1.data(); n = s.tab1.size(); esz = sizeof(double); break;
    default: return EXO_ERR_ARG;
  }
  *count = n;
  if (buf) {
    if (cap < n * esz) return EXO_ERR_ARG;
    std::memcpy(buf, src, n * esz);
  }
  return EXO_OK;
}

size_t exo_footprint(const exo_config *cfg) {
  if (!cfg) return 0;
  Setup s;
  if (!build_setup(*cfg, &s).empty()) return 0;
  return footprint_bytes(s);
}

int exo_conserved_sums(exo_handle *h, double *out5) {
  return guarded(h, [&] {
    if (!out5) throw std::string("null output");
    if (!h->have_state) throw std::string("no state uploaded");
    launch_sums(h->L, h->d_sums);
    HIPCHK(hipMemcpyAsync(out5, h->d_sums, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

}  // extern "C"

namespace {
static_assert(K_NKIND == EXO_NKINDS, "EXO_NKINDS must count the launch kinds");

// tuned widths and shapes per configuration of the instances of a launcher
std::mutex g_tune_mu;
std::map<std::string, std::vector<int>> g_tune_cache;

std::string tune_key(const std::vector<exo_handle *> &hs) {
  std::string k;
  int dev = 0;
  (void)hipGetDevice(&dev);
  k += std::to_string(dev);
  for (const exo_handle *h : hs) {
    const exo_config &c = h->setup.cfg;
    k += "|" + std::to_string(c.nx) + "," + std::to_string(c.nx1) + "," + std::to_string(c.eos) +
         "," + std::to_string(c.rsolver) + "," + std::to_string(c.implicit_flag) + "," +
         std::to_string(c.ntracer) + "," + std::to_string(c.coriolis) + "," +
         std::to_string(c.precision);
  }
  return k;
}

// tune a launcher, or apply the result of an earlier tuning of the same
// configurations
void tune_cached(Launcher &L, const std::vector<exo_handle *> &hs, int reps) {
  std::string key = tune_key(hs);
  std::vector<int> got;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) got = it->second;
  }
  if (got.empty()) {
    tune_launcher(L, reps);
    got.assign(2 * K_NKIND, 0);
    for (int q = 0; q < K_NKIND; ++q) {
      got[q] = L.cap[q];
      got[K_NKIND + q] = L.shape[q];
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = got;
  } else {
    for (int q = 0; q < K_NKIND; ++q) {
      set_shape(L, q, got[K_NKIND + q]);
      L.cap[q] = got[q];
    }
  }
}

void report_tuning(const Launcher &L, int *caps, int ncaps) {
  for (int q = 0; caps && q < ncaps && q < 2 * K_NKIND; ++q)
    caps[q] = q < K_NKIND ? L.cap[q] : L.shape[q - K_NKIND];
}
}  // namespace
extern "C" {

int exo_tune(exo_handle *h, int reps, int *caps, int ncaps) {
  return guarded(h, [&] {
    if (!h->have_state) throw std::string("no state uploaded");
    tune_cached(h->L, std::vector<exo_handle *>{h}, reps > 0 ? reps : 3);
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->exec = nullptr;
    h->graph = nullptr;
    report_tuning(h->L, caps, ncaps);
  });
}

// --- batches ---------------------------------------------------------------
}  // extern "C"

namespace {
void batch_free_graph(exo_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void batch_ready(exo_batch *b) {
  for (exo_handle *m : b->members) {
    if (!m->have_state) throw std::string("a batch member has no state uploaded");
    HIPCHK(hipStreamSynchronize(m->stream));
  }
}
}  // namespace
extern "C" {

int exo_batch_create(exo_handle *const *members, int n, int flags, exo_batch **out) {
  if (!members || n <= 0 || !out) return EXO_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < n; ++i)
    if (!members[i]) return EXO_ERR_ARG;
  exo_batch *b = new exo_batch();
  b->members.assign(members, members + n);
  b->use_graph = flags & EXO_BATCH_GRAPH ? 1 : 0;
  std::vector<void *> allocs;
  try {
    HIPCHK(hipStreamCreateWithFlags(&b->L.stream, hipStreamNonBlocking));
    for (int i = 0; i < n; ++i) HIPCHK(hipStreamSynchronize(members[i]->stream));
    launcher_setup(b->L, b->members, allocs);
    void *p;
    HIPCHK(hipMalloc(&p, (size_t)n * 5 * sizeof(double)));
    allocs.push_back(p);
    b->d_sums = (double *)p;
  } catch (const std::string &e) {
    fprintf(stderr, "exo_batch_create: %s\n", e.c_str());
    for (void *p : allocs) (void)hipFree(p);
    if (b->L.stream) (void)hipStreamDestroy(b->L.stream);
    delete b;
    return EXO_ERR_HIP;
  }
  *out = b;
  return EXO_OK;
}

int exo_batch_destroy(exo_batch *b) {
  if (!b) return EXO_ERR_ARG;
  if (b->L.stream) (void)hipStreamSynchronize(b->L.stream);
  batch_free_graph(b);
  for (hipEvent_t e : b->L.ev_pool) (void)hipEventDestroy(e);
  (void)hipFree(b->L.d_inst);
  (void)hipFree(b->L.d_bstart);
  (void)hipFree(b->d_sums);
  if (b->L.stream) (void)hipStreamDestroy(b->L.stream);
  delete b;
  return EXO_OK;
}

const char *exo_batch_last_error(const exo_batch *b) { return b ? b->err.c_str() : "null batch"; }

int exo_batch_step(exo_batch *b, int nsteps, double dt_fixed) {
  return guarded(b, [&] {
    if (nsteps < 0) throw std::string("nsteps must be >= 0");
    batch_ready(b);
    if (b->use_graph && !b->L.prof) {
      if (!b->exec || b->graph_dt_fixed != dt_fixed) {
        batch_free_graph(b);
        HIPCHK(hipStreamBeginCapture(b->L.stream, hipStreamCaptureModeThreadLocal));
        try {
          run_step(b->L, dt_fixed);
        } catch (...) {
          hipGraph_t gr = nullptr;
          (void)hipStreamEndCapture(b->L.stream, &gr);
          if (gr) (void)hipGraphDestroy(gr);
          throw;
        }
        HIPCHK(hipStreamEndCapture(b->L.stream, &b->graph));
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_dt_fixed = dt_fixed;
      }
      for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(b->exec, b->L.stream));
    } else {
      for (int s = 0; s < nsteps; ++s) run_step(b->L, dt_fixed);
    }
    HIPCHK(hipStreamSynchronize(b->L.stream));
  });
}

int exo_batch_sums(exo_batch *b, double *out) {
  return guarded(b, [&] {
    if (!out) throw std::string("null output");
    batch_ready(b);
    launch_sums(b->L, b->d_sums);
    HIPCHK(hipMemcpyAsync(out, b->d_sums, b->members.size() * 5 * sizeof(double),
                          hipMemcpyDeviceToHost, b->L.stream));
    HIPCHK(hipStreamSynchronize(b->L.stream));
  });
}

int exo_batch_tune(exo_batch *b, int reps, int *caps, int ncaps) {
  return guarded(b, [&] {
    batch_ready(b);
    batch_free_graph(b);
    tune_cached(b->L, b->members, reps > 0 ? reps : 3);
    report_tuning(b->L, caps, ncaps);
  });
}

int exo_batch_set_caps(exo_batch *b, const int *caps, int ncaps) {
  return guarded(b, [&] {
    if (!caps) throw std::string("null caps");
    batch_free_graph(b);
    for (int q = 0; q < ncaps && q < 2 * K_NKIND; ++q) {
      int v = caps[q] > 0 ? caps[q] : 0;
      if (q < K_NKIND) b->L.cap[q] = v;
      else set_shape(b->L, q - K_NKIND, v);
    }
  });
}

int exo_batch_profile_begin(exo_batch *b) {
  return guarded(b, [&] { prof_begin(b->L); });
}

int exo_batch_profile_end(exo_batch *b, double *ms, int *counts, int nkinds) {
  return guarded(b, [&] { prof_end(b->L, ms, counts, nkinds); });
}

int exo_batch_sync(exo_batch *b) {
  return guarded(b, [&] { HIPCHK(hipStreamSynchronize(b->L.stream)); });
}

}  // extern "C"
