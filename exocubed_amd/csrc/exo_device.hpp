// exo_device.hpp -- per-face / per-cell device arithmetic of the cubed-sphere
// hydro update for gfx950.  Every expression keeps the operation order of the
// reference source it restates (cited per function) so that the FP64 build,
// compiled without floating-point contraction, reproduces the reference
// arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace exo {

constexpr int NG = 3;
constexpr double PI = 3.14159265358979323846;
constexpr double TINY_NUMBER = 1.0e-20;
enum { IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IPR = 4, IEN = 4 };
enum { EOS_SW = 0, EOS_IDEAL = 1 };
enum { RS_ROE_SW = 0, RS_LMARS_SW = 1, RS_HLLC = 2, RS_LMARS = 3 };

#define EXO_HD __host__ __device__ __forceinline__

template <typename R>
EXO_HD R sqr(R x) { return x * x; }

// ---------------------------------------------------------------------------
// x / d in FP64 with the steps of gfx950's IEEE division (the compiler's
// v_div_scale / v_rcp / Newton / v_div_fmas / v_div_fixup sequence) minus its
// range scaling: the reciprocal refined by two Newton steps, the quotient, one
// remainder correction (Markstein, IBM J. Res. Dev. 34(1), 1990), then
// v_div_fixup for zeros, infinities and NaNs.  v_div_scale changes neither
// operand, and v_div_fmas is a plain FMA, unless |x| < 2^-969, d is
// subnormal, |d| >= 2^1022, |x / d| < 2^-1022 or x / d is within 2^-256 of
// overflow; outside those ranges (none of which the hydro quantities reach)
// the result is the IEEE quotient bit for bit, for 8 instructions instead of
// 11 and without the VCC hand-off between v_div_scale and v_div_fmas.
EXO_HD double qdiv(double x, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  double q = x * r;
  double rem = fma(-d, q, x);
  return __builtin_amdgcn_div_fixup(fma(rem, r, q), d, x);
#else
  return x / d;
#endif
}
EXO_HD float qdiv(float x, float d) { return x / d; }

// ---------------------------------------------------------------------------
// Correctly rounded quotient x / d from y = RN(1/d), one IEEE division shared
// by every quotient with the same divisor.  q0 = RN(x y) lies within one ulp
// of x/d, so the remainder x - q0 d is exact under FMA and RN(q0 + r y) is the
// correctly rounded x/d (Markstein's theorem, IBM J. Res. Dev. 34(1), 1990;
// the basis of the IA-64 division sequences).  It needs x/d, q0 and r away
// from overflow and the subnormal range, which holds for the geometric and
// thermodynamic quantities divided here.  Three instructions instead of the
// ~10 of an IEEE division, with the identical result (one exception: x = -0
// gives +0, an equal value, where IEEE division keeps the sign).
EXO_HD double div_rn(double x, double d, double y) {
  double q0 = x * y;
  double r = fma(-q0, d, x);
  return fma(r, y, q0);
}

// ---------------------------------------------------------------------------
// Deterministic exp / log / pow (FP64) used by the hydrostatic decomposition
// (change_to_perturbation.cpp:63-65,109-115,181-182; decomposition.cpp:357-360).
// Cody-Waite reduction with a degree-13 Taylor polynomial for exp, and the
// atanh series of (m-1)/(m+1) for log; the same method runs on the host side
// of the tests, so device and host agree to the last bit.
constexpr double LN2_HI = 6.93147180369123816490e-01;
constexpr double LN2_LO = 1.90821492927058770002e-10;
constexpr double INV_LN2 = 1.44269504088896338700e+00;

EXO_HD double dexp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  double kd = floor(x * INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 6227020800.0;
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  return ldexp(p, k);
}

EXO_HD double dlog(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int k;
  double m = frexp(x, &k);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    k = k - 1;
  }
  double f = qdiv(m - 1.0, m + 1.0);
  double s = f * f;
  double p = 1.0 / 25.0;
  p = 1.0 / 23.0 + s * p;
  p = 1.0 / 21.0 + s * p;
  p = 1.0 / 19.0 + s * p;
  p = 1.0 / 17.0 + s * p;
  p = 1.0 / 15.0 + s * p;
  p = 1.0 / 13.0 + s * p;
  p = 1.0 / 11.0 + s * p;
  p = 1.0 / 9.0 + s * p;
  p = 1.0 / 7.0 + s * p;
  p = 1.0 / 5.0 + s * p;
  p = 1.0 / 3.0 + s * p;
  double lm = 2.0 * f + (2.0 * f) * (s * p);
  double kd = (double)k;
  return kd * LN2_HI + (lm + kd * LN2_LO);
}

EXO_HD double dpow(double x, double y) { return dexp(y * dlog(x)); }

// ---------------------------------------------------------------------------
// WENO5 value at the face between b and c seen from cell c: Jiang-Shu
// smoothness indicators, eps 1e-6, linear weights 3/10, 6/10, 1/10 with the
// 3/10 term summed first (tests/test_weno.cpp:20-29 returns 2.5000000000000004
// for 1..5 only in that order).  Reconstruction::Weno5X* apply it per face.
template <typename R>
EXO_HD R weno5(R phim2, R phim1, R phi, R phip1, R phip2) {
  R p0 = (R(1.) / R(3.)) * phi + (R(5.) / R(6.)) * phim1 - (R(1.) / R(6.)) * phim2;
  R p1 = (R(-1.) / R(6.)) * phip1 + (R(5.) / R(6.)) * phi + (R(1.) / R(3.)) * phim1;
  R p2 = (R(1.) / R(3.)) * phip2 - (R(7.) / R(6.)) * phip1 + (R(11.) / R(6.)) * phi;
  R b0 = R(13.) / R(12.) * sqr(phi - R(2.) * phim1 + phim2) +
         R(.25) * sqr(R(3.) * phi - R(4.) * phim1 + phim2);
  R b1 = R(13.) / R(12.) * sqr(phip1 - R(2.) * phi + phim1) + R(.25) * sqr(phip1 - phim1);
  R b2 = R(13.) / R(12.) * sqr(phip2 - R(2.) * phip1 + phi) +
         R(.25) * sqr(phip2 - R(4.) * phip1 + R(3.) * phi);
  R a0 = qdiv(R(.3), sqr(b0 + R(1e-6)));
  R a1 = qdiv(R(.6), sqr(b1 + R(1e-6)));
  R a2 = qdiv(R(.1), sqr(b2 + R(1e-6)));
  return qdiv(a0 * p0 + a1 * p1 + a2 * p2, a0 + a1 + a2);
}

// ---------------------------------------------------------------------------
// panel <-> lat-lon vector maps, src/exo3/cubed_sphere_transform.cpp:11-61
EXO_HD void vec_abp_from_rll(double X, double Y, int blockID, double U, double V,
                             double *V2, double *V3) {
  double C = sqrt(1 + X * X);
  double D = sqrt(1 + Y * Y);
  double E = sqrt(X * X + Y * Y);
  double delta = sqrt(1 + X * X + Y * Y);
  switch (blockID) {
    case 2: case 3: case 4: case 6:
      *V2 = X * Y / delta * U - V;
      *V3 = C * D / delta * U;
      break;
    case 1:
      *V2 = -D * Y / (delta * E) * U - D * X / E * V;
      *V3 = C * X / (delta * E) * U - C * Y / E * V;
      break;
    default:
      *V2 = D * Y / (delta * E) * U + D * X / E * V;
      *V3 = -C * X / (delta * E) * U + C * Y / E * V;
      break;
  }
}

EXO_HD void vec_rll_from_abp(double X, double Y, int blockID, double V2, double V3,
                             double *U, double *V) {
  double C = sqrt(1 + X * X);
  double D = sqrt(1 + Y * Y);
  double E = sqrt(X * X + Y * Y);
  double delta = sqrt(1 + X * X + Y * Y);
  switch (blockID) {
    case 2: case 3: case 4: case 6:
      *U = delta / (C * D) * V3;
      *V = X * Y / (C * D) * V3 - V2;
      break;
    case 1:
      *U = -delta * Y / (D * E) * V2 + delta * X / (C * E) * V3;
      *V = -X / (D * E) * V2 - Y / (C * E) * V3;
      break;
    default:
      *U = delta * Y / (D * E) * V2 - delta * X / (C * E) * V3;
      *V = X / (D * E) * V2 + Y / (C * E) * V3;
      break;
  }
}

// ---------------------------------------------------------------------------
// Riemann fluxes of one face.  q[] holds (IDN, IVX, IVY, IVZ[, IPR]) left and
// right primitives in the panel frame; f[] receives the global-frame,
// covariant-projected flux.  cosf/sinf: face metric (Face2/3Metric,
// gnomonic_equiangle.cpp:494-569); cth: the solver's covariant factor of the
// row (e.g. roe_shallow_yz.cpp:97-110); cosc: cell cosine (HLLC x1 kinetic
// energy, hllc_transform.cpp:95-106).
template <typename R>
EXO_HD void local_frame(int ivx, R cosf, R sinf, R *q) {
  // PrimToLocal2/3, gnomonic_equiangle.cpp:604-718
  R g23 = cosf, gi = qdiv(R(1.0), sinf * sinf);
  R uu1 = q[IVX], uu2 = q[IVY], uu3 = q[IVZ];
  if (ivx == IVY) {
    R T22 = qdiv(R(1.0), R(sqrt(gi))), T32 = g23 / sqrt(R(1.0)), T33 = sqrt(R(1.0));
    q[IVX] = R(1.0) * uu1;
    q[IVY] = T22 * uu2;
    q[IVZ] = T32 * uu2 + T33 * uu3;
  } else {
    R T22 = sqrt(R(1.0)), T23 = g23 / sqrt(R(1.0)), T33 = qdiv(R(1.0), R(sqrt(gi)));
    q[IVX] = R(1.0) * uu1;
    q[IVY] = T22 * uu2 + T23 * uu3;
    q[IVZ] = T33 * uu3;
  }
}

template <typename R>
EXO_HD void flux_to_global(int ivx, R cosf, R sinf, R *f) {
  // FluxToGlobal2/3, gnomonic_equiangle.cpp:752-832
  R g23 = cosf, gi = qdiv(R(1.0), sinf * sinf);
  R txx = f[IVX], txy = f[IVY], txz = f[IVZ];
  if (ivx == IVY) {
    R T22 = sqrt(gi), T32 = -sqrt(gi) * g23 / R(1.0), T33 = R(1.0) / sqrt(R(1.0));
    f[IVX] = R(1.0) * txx;
    f[IVY] = T22 * txy;
    f[IVZ] = T32 * txy + T33 * txz;
  } else {
    R T22 = R(1.0) / sqrt(R(1.0)), T23 = -g23 / R(1.0) * sqrt(gi), T33 = sqrt(gi);
    f[IVX] = R(1.0) * txx;
    f[IVY] = T22 * txy + T23 * txz;
    f[IVZ] = T33 * txz;
  }
}

// roe_shallow_yz.cpp:39-87
template <typename R>
EXO_HD void roe_sw(int ivx, const R *wli, const R *wri, R *fl) {
  int ivy = ivx == IVY ? IVZ : IVY;
  R ubar = qdiv(wli[ivx] * sqrt(wli[IDN]) + wri[ivx] * sqrt(wri[IDN]),
               sqrt(wli[IDN]) + sqrt(wri[IDN]));
  R vbar = qdiv(wli[ivy] * sqrt(wli[IDN]) + wri[ivy] * sqrt(wri[IDN]),
               sqrt(wli[IDN]) + sqrt(wri[IDN]));
  R cbar = sqrt(R(0.5) * (wli[IDN] + wri[IDN]));
  R delh = wri[IDN] - wli[IDN];
  R delu = wri[ivx] - wli[ivx];
  R delv = wri[ivy] - wli[ivy];
  R hbar = sqrt(wli[IDN] * wri[IDN]);
  R a1 = qdiv(R(0.5) * (cbar * delh - hbar * delu), cbar);
  R a2 = hbar * delv;
  R a3 = qdiv(R(0.5) * (cbar * delh + hbar * delu), cbar);
  R w00 = a1, w01 = a1 * (ubar - cbar), w02 = a1 * vbar;
  R w12 = a2;
  R w20 = a3, w21 = a3 * (ubar + cbar), w22 = a3 * vbar;
  R s0 = fabs(ubar - cbar), s1 = fabs(ubar), s2 = fabs(ubar + cbar);
  R fd = R(0.5) * (wli[IDN] * wli[ivx] + wri[IDN] * wri[ivx]);
  R fx = R(0.5) * (wli[IDN] * wli[ivx] * wli[ivx] + R(0.5) * wli[IDN] * wli[IDN] +
                   wri[IDN] * wri[ivx] * wri[ivx] + R(0.5) * wri[IDN] * wri[IDN]);
  R fy = R(0.5) * (wli[IDN] * wli[ivx] * wli[ivy] + wri[IDN] * wri[ivx] * wri[ivy]);
  // wave r = 0, 1, 2 in order (row 1 of the wave matrix is 0, 0, a2)
  fd -= R(0.5) * s0 * w00;
  fx -= R(0.5) * s0 * w01;
  fy -= R(0.5) * s0 * w02;
  fd -= R(0.5) * s1 * R(0.);
  fx -= R(0.5) * s1 * R(0.);
  fy -= R(0.5) * s1 * w12;
  fd -= R(0.5) * s2 * w20;
  fx -= R(0.5) * s2 * w21;
  fy -= R(0.5) * s2 * w22;
  fl[IDN] = fd;
  fl[IVX] = R(0.);
  fl[ivx] = fx;
  fl[ivy] = fy;
  fl[IVX] = R(0.);
}

// lmars_shallow_yz.cpp:44-80
template <typename R>
EXO_HD void lmars_sw(int ivx, const R *wli, const R *wri, R *fl) {
  int ivy = ivx == IVY ? IVZ : IVY;
  R cbar = sqrt(R(0.5) * (wli[IDN] + wri[IDN]));
  R hbar = R(0.5) * (wli[IDN] + wri[IDN]) + R(0.5) * cbar * (wli[ivx] - wri[ivx]);
  R ubar = R(0.5) * (wli[ivx] + wri[ivx]) + qdiv(R(0.5), cbar) * (wli[IDN] - wri[IDN]);
  if (ubar > R(0.)) {
    fl[IDN] = ubar * wli[IDN];
    fl[ivx] = ubar * wli[IDN] * wli[ivx] + R(0.5) * hbar * hbar;
    fl[ivy] = ubar * wli[IDN] * wli[ivy];
  } else {
    fl[IDN] = ubar * wri[IDN];
    fl[ivx] = ubar * wri[IDN] * wri[ivx] + R(0.5) * hbar * hbar;
    fl[ivy] = ubar * wri[IDN] * wri[ivy];
  }
  fl[IVX] = R(0.);
}

// hllc_transform.cpp:54-223 (single face)
template <typename R>
EXO_HD void hllc(int ivx, R gamma, const R *wli, const R *wri, R *fl, R cosc) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  R igm1 = qdiv(R(1.0), gamma - R(1.0));
  R wl[5] = {wli[IDN], wli[ivx], wli[ivy], wli[ivz], wli[IPR]};
  R wr[5] = {wri[IDN], wri[ivx], wri[ivy], wri[ivz], wri[IPR]};
  R cl = sqrt(qdiv((R(1.) + (gamma - R(1)) * R(1.) / R(1.)) * wl[IPR], wl[IDN]));
  R cr = sqrt(qdiv((R(1.) + (gamma - R(1)) * R(1.) / R(1.)) * wr[IPR], wr[IDN]));
  R KE_l, KE_r;
  if (ivx == IVX) {
    R vy = wl[IVY] + wl[IVZ] * cosc, vz = wl[IVZ] + wl[IVY] * cosc;
    KE_l = R(0.5) * wl[IDN] * (sqr(wl[IVX]) + wl[IVY] * vy + wl[IVZ] * vz);
    vy = wr[IVY] + wr[IVZ] * cosc;
    vz = wr[IVZ] + wr[IVY] * cosc;
    KE_r = R(0.5) * wr[IDN] * (sqr(wr[IVX]) + wr[IVY] * vy + wr[IVZ] * vz);
  } else {
    KE_l = R(0.5) * wl[IDN] * (sqr(wl[IVX]) + sqr(wl[IVY]) + sqr(wl[IVZ]));
    KE_r = R(0.5) * wr[IDN] * (sqr(wr[IVX]) + sqr(wr[IVY]) + sqr(wr[IVZ]));
  }
  R el = wl[IPR] * igm1 + KE_l;
  R er = wr[IPR] * igm1 + KE_r;
  R rhoa = R(.5) * (wl[IDN] + wr[IDN]);
  R ca = R(.5) * (cl + cr);
  R pmid = R(.5) * (wl[IPR] + wr[IPR] + (wl[IVX] - wr[IVX]) * rhoa * ca);
  R ql = (pmid <= wl[IPR])
             ? R(1.0)
             : sqrt(R(1.0) + qdiv(gamma + R(1), R(2) * gamma) * (qdiv(pmid, wl[IPR]) - R(1.0)));
  R qr = (pmid <= wr[IPR])
             ? R(1.0)
             : sqrt(R(1.0) + qdiv(gamma + R(1), R(2) * gamma) * (qdiv(pmid, wr[IPR]) - R(1.0)));
  R al = wl[IVX] - cl * ql;
  R ar = wr[IVX] + cr * qr;
  R bp = ar > R(0.0) ? ar : R(TINY_NUMBER);
  R bm = al < R(0.0) ? al : -R(TINY_NUMBER);
  R vxl = wl[IVX] - al, vxr = wr[IVX] - ar;
  R tl = wl[IPR] + vxl * wl[IDN] * wl[IVX];
  R tr = wr[IPR] + vxr * wr[IDN] * wr[IVX];
  R ml = wl[IDN] * vxl;
  R mr = -(wr[IDN] * vxr);
  R am = qdiv(tl - tr, ml + mr);
  R cp = qdiv(ml * tr + mr * tl, ml + mr);
  cp = cp > R(0.0) ? cp : R(0.0);
  vxl = wl[IVX] - bm;
  vxr = wr[IVX] - bp;
  R fLd = wl[IDN] * vxl, fRd = wr[IDN] * vxr;
  R fLx = wl[IDN] * wl[IVX] * vxl + wl[IPR], fRx = wr[IDN] * wr[IVX] * vxr + wr[IPR];
  R fLy = wl[IDN] * wl[IVY] * vxl, fRy = wr[IDN] * wr[IVY] * vxr;
  R fLz = wl[IDN] * wl[IVZ] * vxl, fRz = wr[IDN] * wr[IVZ] * vxr;
  R fLe = el * vxl + wl[IPR] * wl[IVX], fRe = er * vxr + wr[IPR] * wr[IVX];
  R sl, sr, sm;
  if (am >= R(0.0)) {
    sl = qdiv(am, am - bm);
    sr = R(0.0);
    sm = qdiv(-bm, am - bm);
  } else {
    sl = R(0.0);
    sr = qdiv(-am, bp - am);
    sm = qdiv(bp, bp - am);
  }
  fl[IDN] = sl * fLd + sr * fRd;
  fl[ivx] = sl * fLx + sr * fRx + sm * cp;
  fl[ivy] = sl * fLy + sr * fRy;
  fl[ivz] = sl * fLz + sr * fRz;
  fl[IEN] = sl * fLe + sr * fRe + sm * cp * am;
}

// lmars.cpp:302-362, NVAPOR = NCLOUD = 0 (kappa = 1/(gamma-1))
template <typename R>
EXO_HD void lmars(int ivx, R gamma, const R *wli, const R *wri, R *fl) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  R kappal = qdiv(R(1.), gamma - R(1.)) * R(1.) / R(1.);
  R kappar = qdiv(R(1.), gamma - R(1.)) * R(1.) / R(1.);
  R hl = qdiv(wli[IPR], wli[IDN]) * (kappal + R(1.)) +
         R(0.5) * (sqr(wli[ivx]) + sqr(wli[ivy]) + sqr(wli[ivz]));
  R hr = qdiv(wri[IPR], wri[IDN]) * (kappar + R(1.)) +
         R(0.5) * (sqr(wri[ivx]) + sqr(wri[ivy]) + sqr(wri[ivz]));
  R rhobar = R(0.5) * (wli[IDN] + wri[IDN]);
  R cbar = sqrt(qdiv(R(0.5) * (R(1.) + (qdiv(R(1.), kappar) + qdiv(R(1.), kappal)) / R(2.)) *
                         (wli[IPR] + wri[IPR]),
                     rhobar));
  R pbar = R(0.5) * (wli[IPR] + wri[IPR]) + R(0.5) * (rhobar * cbar) * (wli[ivx] - wri[ivx]);
  R ubar =
      R(0.5) * (wli[ivx] + wri[ivx]) + qdiv(R(0.5), rhobar * cbar) * (wli[IPR] - wri[IPR]);
  R rdl = R(1.), rdr = R(1.);
  if (ubar > R(0.)) {
    fl[IDN] = ubar * wli[IDN] * rdl;
    fl[ivx] = ubar * wli[IDN] * wli[ivx] + pbar;
    fl[ivy] = ubar * wli[IDN] * wli[ivy];
    fl[ivz] = ubar * wli[IDN] * wli[ivz];
    fl[IEN] = ubar * wli[IDN] * hl;
  } else {
    fl[IDN] = ubar * wri[IDN] * rdr;
    fl[ivx] = ubar * wri[IDN] * wri[ivx] + pbar;
    fl[ivy] = ubar * wri[IDN] * wri[ivy];
    fl[ivz] = ubar * wri[IDN] * wri[ivz];
    fl[IEN] = ubar * wri[IDN] * hr;
  }
}

// Hydro::RiemannSolver for one face: frame transforms around the solver as
// each solver file does them.  wl / wr are modified (local frame).  The
// configuration pairs shallow water with the Roe or LMARS shallow-water
// solver and the ideal gas with HLLC or LMARS, so a caller that fixes eos at
// compile time compiles only its two solvers.
template <typename R>
EXO_HD void riemann(int eos, int rs, int ivx, R gamma, R *wl, R *wr, R *f, R cth, R cosf,
                    R sinf, R cosc) {
  f[0] = f[1] = f[2] = f[3] = R(0.);
  if (eos == EOS_IDEAL) f[4] = R(0.);
  bool maps = eos == EOS_SW || rs != RS_LMARS;  // lmars.cpp maps nothing
  if (ivx != IVX && maps) {
    local_frame(ivx, cosf, sinf, wl);
    local_frame(ivx, cosf, sinf, wr);
  }
  if (eos == EOS_SW) {
    if (rs == RS_ROE_SW) roe_sw(ivx, wl, wr, f);
    else lmars_sw(ivx, wl, wr, f);
  } else if (rs == RS_HLLC) {
    hllc(ivx, gamma, wl, wr, f, cosc);
  } else {
    lmars(ivx, gamma, wl, wr, f);
  }
  if (maps) {
    if (ivx != IVX) flux_to_global(ivx, cosf, sinf, f);
    R ty = f[IVY], tz = f[IVZ];
    f[IVY] = ty + tz * cth;
    f[IVZ] = tz + ty * cth;
  }
}


// ---------------------------------------------------------------------------
// 5x5 algebra of the implicit column solve (Eigen in full_correction.cpp):
// k-ordered sums and LU with partial pivoting, the same as the oracle.  Sums
// of products are fused multiply-add chains (t = fma(a_k, b_k, t) for k = 1..4
// after t = a_0 b_0), as Eigen's fixed-size products and its LU rank-1 update
// are when built for FMA hardware; the oracle uses the same chains.
struct M5 {
  double a[5][5];
};
struct V5 {
  double a[5];
};

EXO_HD M5 m5_mul(const M5 &A, const M5 &B) {
  M5 C;
#pragma unroll
  for (int r = 0; r < 5; ++r)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double t = A.a[r][0] * B.a[0][c];
#pragma unroll
      for (int k = 1; k < 5; ++k) t = fma(A.a[r][k], B.a[k][c], t);
      C.a[r][c] = t;
    }
  return C;
}

EXO_HD V5 m5_mulv(const M5 &A, const V5 &x) {
  V5 y;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    double t = A.a[r][0] * x.a[0];
#pragma unroll
    for (int k = 1; k < 5; ++k) t = fma(A.a[r][k], x.a[k], t);
    y.a[r] = t;
  }
  return y;
}

EXO_HD M5 m5_inv(const M5 &A) {
  double lu[5][5], rd[5];  // rd: reciprocals of the pivots
  int piv[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) {
#pragma unroll
    for (int c = 0; c < 5; ++c) lu[r][c] = A.a[r][c];
    piv[r] = r;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    int p = c;
    double best = fabs(lu[c][c]);
#pragma unroll
    for (int r = c + 1; r < 5; ++r)
      if (fabs(lu[r][c]) > best) {
        best = fabs(lu[r][c]);
        p = r;
      }
#pragma unroll
    for (int r = c + 1; r < 5; ++r) {
      if (p == r) {
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          double t = lu[c][q];
          lu[c][q] = lu[r][q];
          lu[r][q] = t;
        }
        int t = piv[c];
        piv[c] = piv[r];
        piv[r] = t;
      }
    }
    double yc = qdiv(1.0, lu[c][c]);
    rd[c] = yc;
#pragma unroll
    for (int r = c + 1; r < 5; ++r) {
      double l = div_rn(lu[r][c], lu[c][c], yc);
      lu[r][c] = l;
#pragma unroll
      for (int q = c + 1; q < 5; ++q) lu[r][q] = fma(-l, lu[c][q], lu[r][q]);
    }
  }
  M5 X;
#pragma unroll
  for (int col = 0; col < 5; ++col) {
    double y[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      double t = piv[r] == col ? 1.0 : 0.0;
#pragma unroll
      for (int q = 0; q < r; ++q) t = fma(-lu[r][q], y[q], t);
      y[r] = t;
    }
#pragma unroll
    for (int r = 4; r >= 0; --r) {
      double t = y[r];
#pragma unroll
      for (int q = r + 1; q < 5; ++q) t = fma(-lu[r][q], X.a[q][col], t);
      X.a[r][col] = div_rn(t, lu[r][r], rd[r]);
    }
  }
  return X;
}

// The same algebra on the 3x3 blocks (density, x1 momentum, energy) of the
// partial column correction (partial_correction.cpp).
struct M3 {
  double a[3][3];
};
struct V3 {
  double a[3];
};

EXO_HD M3 m3_mul(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double t = A.a[r][0] * B.a[0][c];
#pragma unroll
      for (int k = 1; k < 3; ++k) t = fma(A.a[r][k], B.a[k][c], t);
      C.a[r][c] = t;
    }
  return C;
}

EXO_HD V3 m3_mulv(const M3 &A, const V3 &x) {
  V3 y;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double t = A.a[r][0] * x.a[0];
#pragma unroll
    for (int k = 1; k < 3; ++k) t = fma(A.a[r][k], x.a[k], t);
    y.a[r] = t;
  }
  return y;
}

EXO_HD M3 m3_inv(const M3 &A) {
  double lu[3][3], rd[3];
  int piv[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) lu[r][c] = A.a[r][c];
    piv[r] = r;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int p = c;
    double best = fabs(lu[c][c]);
#pragma unroll
    for (int r = c + 1; r < 3; ++r)
      if (fabs(lu[r][c]) > best) {
        best = fabs(lu[r][c]);
        p = r;
      }
#pragma unroll
    for (int r = c + 1; r < 3; ++r) {
      if (p == r) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double t = lu[c][q];
          lu[c][q] = lu[r][q];
          lu[r][q] = t;
        }
        int t = piv[c];
        piv[c] = piv[r];
        piv[r] = t;
      }
    }
    double yc = qdiv(1.0, lu[c][c]);
    rd[c] = yc;
#pragma unroll
    for (int r = c + 1; r < 3; ++r) {
      double l = div_rn(lu[r][c], lu[c][c], yc);
      lu[r][c] = l;
#pragma unroll
      for (int q = c + 1; q < 3; ++q) lu[r][q] = fma(-l, lu[c][q], lu[r][q]);
    }
  }
  M3 X;
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    double y[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double t = piv[r] == col ? 1.0 : 0.0;
#pragma unroll
      for (int q = 0; q < r; ++q) t = fma(-lu[r][q], y[q], t);
      y[r] = t;
    }
#pragma unroll
    for (int r = 2; r >= 0; --r) {
      double t = y[r];
#pragma unroll
      for (int q = r + 1; q < 3; ++q) t = fma(-lu[r][q], X.a[q][col], t);
      X.a[r][col] = div_rn(t, lu[r][r], rd[r]);
    }
  }
  return X;
}

// flux_decomposition.hpp:34-61 RoeAverage (NVAPOR = 0)
EXO_HD void roe_average(double *prim, double gm1, const double *wl, const double *wr) {
  double sqrtdl = sqrt(wl[IDN]), sqrtdr = sqrt(wr[IDN]);
  double isdlpdr = qdiv(1.0, sqrtdl + sqrtdr);
  prim[IDN] = sqrtdl * sqrtdr;
  prim[IVX] = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  prim[IVY] = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  prim[IVZ] = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  double el = qdiv(wl[IPR], gm1) + 0.5 * wl[IDN] * (sqr(wl[IVX]) + sqr(wl[IVY]) + sqr(wl[IVZ]));
  double er = qdiv(wr[IPR], gm1) + 0.5 * wr[IDN] * (sqr(wr[IVX]) + sqr(wr[IVY]) + sqr(wr[IVZ]));
  double hbar = (qdiv(el + wl[IPR], sqrtdl) + qdiv(er + wr[IPR], sqrtdr)) * isdlpdr;
  prim[IPR] = qdiv((hbar - 0.5 * (sqr(prim[IVX]) + sqr(prim[IVY]) + sqr(prim[IVZ]))) * gm1,
                   gm1 + 1.) *
              prim[IDN];
}

// R |Lambda| R^-1 from Eigenvalue / Eigenvector (flux_decomposition.hpp:63-96).
// The product runs over the structurally non-zero terms only: R (|Lambda|
// scales its columns) has zeros at (0,3) (0,4) (1,3) (1,4) (2,4) (3,3), and
// rows 3 and 4 of R^-1 are (-v, 0, 1, 0, 0) and (-w, 0, 0, 1, 0).  A skipped
// term is an exact zero, so for finite entries each sum equals the dense one.
EXO_HD M5 abs_a(const double *prim, double cs, double gm1) {
  double r = prim[IDN], u = prim[IVX], v = prim[IVY], w = prim[IVZ], p = prim[IPR];
  double ke = 0.5 * (u * u + v * v + w * w);
  double hp = qdiv(qdiv(gm1 + 1., gm1) * p, r);
  double h = hp + ke;
  double l0 = fabs(u - cs), l1 = fabs(u), l2 = fabs(u + cs), l3 = fabs(u), l4 = fabs(u);
  // columns 0..2 of R |Lambda|
  double RL[5][3] = {{1. * l0, 1. * l1, 1. * l2},
                     {(u - cs) * l0, u * l1, (u + cs) * l2},
                     {v * l0, v * l1, v * l2},
                     {w * l0, w * l1, w * l2},
                     {(h - u * cs) * l0, ke * l1, (h + u * cs) * l2}};
  // columns 3 and 4 of R |Lambda|: (R(2,3), R(4,3)) = (1, v), (R(3,4), R(4,4)) = (1, w)
  double r23 = 1. * l3, r43 = v * l3, r34 = 1. * l4, r44 = w * l4;
  double d1 = 2. * cs * hp, d2 = 2. * hp;
  double y1 = qdiv(1., d1), y2 = qdiv(1., d2), y3 = qdiv(1., hp);
  double Ri[3][5] = {{div_rn(cs * ke + u * hp, d1, y1), div_rn(-hp - cs * u, d1, y1),
                      div_rn(-v, d2, y2), div_rn(-w, d2, y2), y2},
                     {div_rn(hp - ke, hp, y3), div_rn(u, hp, y3), div_rn(v, hp, y3),
                      div_rn(w, hp, y3), div_rn(-1., hp, y3)},
                     {div_rn(cs * ke - u * hp, d1, y1), div_rn(hp - cs * u, d1, y1),
                      div_rn(-v, d2, y2), div_rn(-w, d2, y2), y2}};
  M5 A;
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double t = RL[a][0] * Ri[0][c];
      t = fma(RL[a][1], Ri[1][c], t);
      A.a[a][c] = fma(RL[a][2], Ri[2][c], t);
    }
  // row 3 of R^-1 enters through R(.,3), row 4 through R(.,4)
  A.a[2][0] = fma(r23, -v, A.a[2][0]);
  A.a[2][2] = fma(r23, 1., A.a[2][2]);
  A.a[4][0] = fma(r43, -v, A.a[4][0]);
  A.a[4][2] = fma(r43, 1., A.a[4][2]);
  A.a[3][0] = fma(r34, -w, A.a[3][0]);
  A.a[3][3] = fma(r34, 1., A.a[3][3]);
  A.a[4][0] = fma(r44, -w, A.a[4][0]);
  A.a[4][3] = fma(r44, 1., A.a[4][3]);
  return A;
}

// FluxJacobian, flux_decomposition.hpp:102-120 (direction x1)
EXO_HD M5 flux_jacobian(double gm1, const double *w) {
  double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ], rho = w[IDN], pres = w[IPR];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * pres / rho) * v1;
  double c2 = (gm1 + 1) / gm1 * pres / rho + s2 / 2 - gm1 * v1 * v1;
  M5 J = {{{0, 1., 0., 0., 0.},
           {gm1 * s2 / 2 - v1 * v1, (2. - gm1) * v1, -gm1 * v2, -gm1 * v3, gm1},
           {-v1 * v2, v2, v1, 0., 0.},
           {-v1 * v3, v3, 0., v1, 0.},
           {c1, c2, -gm1 * v2 * v1, -gm1 * v3 * v1, (gm1 + 1) * v1}}};
  return J;
}

}  // namespace exo
