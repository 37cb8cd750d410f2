// exo_mixed.hpp -- FP32 face fluxes in perturbation form for the mixed-precision
// path.  The face states carry the pressure (shallow water: the geopotential)
// as a perturbation p' about a reference P that is constant along the sweep
// (a per-level reference for x2 / x3 faces, the FP64 hydrostatic face pressure
// psf of the decomposition for x1 faces).  The solvers below return the flux
// with the reference-pressure part of the normal momentum flux removed; the
// FP64 update kernel adds P times the face's metric factors back, so the
// large hydrostatic and metric pressure balances never pass through FP32.
// The algebra follows the FP64 solvers in exo_device.hpp (and the reference
// files they cite) with the pressure split out exactly:
//   LMARS:  pbar = P + pbar',  pbar' = (p'_l + p'_r)/2 + rho c (u_l - u_r)/2
//   HLLC:   t_l = P + t_l', cp = P + cp', sl + sr + sm = 1
// Shallow water instead takes each face about its left cell's FP64 state
// (roe_sw_delta, lmars_sw_delta below).
// Unlike the FP64 path, the FP32 WENO5 uses FMAs: the mixed form is held to
// its error bound, not to the reference's roundings.
#pragma once

#include "exo_device.hpp"

namespace exo {

EXO_HD float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// WENO5 of exo_device.hpp in FP32 with hardware reciprocals for the weights,
// its products and sums written as explicit FMAs.  Explicit, because the edge
// kernel and the flux kernels reconstruct the same panel-edge value and must
// agree bit for bit (the two panels' edge fluxes then agree and mass is
// conserved); compiler contraction could fuse differently in each kernel.
EXO_HD float weno5f(float phim2, float phim1, float phi, float phip1, float phip2) {
  float p0 = fmaf(1.f / 3.f, phi, fmaf(5.f / 6.f, phim1, (-1.f / 6.f) * phim2));
  float p1 = fmaf(-1.f / 6.f, phip1, fmaf(5.f / 6.f, phi, (1.f / 3.f) * phim1));
  float p2 = fmaf(1.f / 3.f, phip2, fmaf(-7.f / 6.f, phip1, (11.f / 6.f) * phi));
  float t0 = fmaf(-2.f, phim1, phi + phim2), s0 = fmaf(3.f, phi, fmaf(-4.f, phim1, phim2));
  float t1 = fmaf(-2.f, phi, phip1 + phim1), s1 = phip1 - phim1;
  float t2 = fmaf(-2.f, phip1, phip2 + phi), s2 = fmaf(3.f, phi, fmaf(-4.f, phip1, phip2));
  float b0 = fmaf((13.f / 12.f) * t0, t0, fmaf(.25f * s0, s0, 1e-6f));
  float b1 = fmaf((13.f / 12.f) * t1, t1, fmaf(.25f * s1, s1, 1e-6f));
  float b2 = fmaf((13.f / 12.f) * t2, t2, fmaf(.25f * s2, s2, 1e-6f));
  float a0 = .3f * frcp(b0 * b0);
  float a1 = .6f * frcp(b1 * b1);
  float a2 = .1f * frcp(b2 * b2);
  return fmaf(a0, p0, fmaf(a1, p1, a2 * p2)) * frcp(a0 + a1 + a2);
}

// q[IDN] full density (geopotential for shallow water), q[IVX..IVZ]
// velocities, q[IPR] pressure perturbation p'; P the reference pressure.
EXO_HD void lmars_pert(int ivx, float gamma, const float *wl, const float *wr, float P,
                       float *fl) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  float kappa = 1.f / (gamma - 1.f);
  float pl = P + wl[IPR], pr = P + wr[IPR];
  float hl = pl * frcp(wl[IDN]) * (kappa + 1.f) +
             0.5f * (wl[ivx] * wl[ivx] + wl[ivy] * wl[ivy] + wl[ivz] * wl[ivz]);
  float hr = pr * frcp(wr[IDN]) * (kappa + 1.f) +
             0.5f * (wr[ivx] * wr[ivx] + wr[ivy] * wr[ivy] + wr[ivz] * wr[ivz]);
  float rhobar = 0.5f * (wl[IDN] + wr[IDN]);
  float cbar = sqrtf(0.5f * (1.f + 1.f / kappa) * (pl + pr) * frcp(rhobar));
  float rc = rhobar * cbar;
  float pbar = 0.5f * (wl[IPR] + wr[IPR]) + 0.5f * rc * (wl[ivx] - wr[ivx]);
  float ubar = 0.5f * (wl[ivx] + wr[ivx]) + 0.5f * frcp(rc) * (wl[IPR] - wr[IPR]);
  const float *w = ubar > 0.f ? wl : wr;
  float h = ubar > 0.f ? hl : hr;
  float m = ubar * w[IDN];
  fl[IDN] = m;
  fl[ivx] = m * w[ivx] + pbar;
  fl[ivy] = m * w[ivy];
  fl[ivz] = m * w[ivz];
  fl[IEN] = m * h;
}

EXO_HD void hllc_pert(int ivx, float gamma, const float *wli, const float *wri, float P,
                      float *fl, float cosc) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  float igm1 = 1.f / (gamma - 1.f);
  float dl = wli[IDN], ul = wli[ivx], vl = wli[ivy], zl = wli[ivz], ql_ = wli[IPR];
  float dr = wri[IDN], ur = wri[ivx], vr = wri[ivy], zr = wri[ivz], qr_ = wri[IPR];
  float pl = P + ql_, pr = P + qr_;
  float cl = sqrtf(gamma * pl * frcp(dl)), cr = sqrtf(gamma * pr * frcp(dr));
  float KE_l, KE_r;
  if (ivx == IVX) {
    float vy = vl + zl * cosc, vz = zl + vl * cosc;
    KE_l = 0.5f * dl * (ul * ul + vl * vy + zl * vz);
    vy = vr + zr * cosc;
    vz = zr + vr * cosc;
    KE_r = 0.5f * dr * (ur * ur + vr * vy + zr * vz);
  } else {
    KE_l = 0.5f * dl * (ul * ul + vl * vl + zl * zl);
    KE_r = 0.5f * dr * (ur * ur + vr * vr + zr * zr);
  }
  float el = pl * igm1 + KE_l, er = pr * igm1 + KE_r;
  float rhoa = .5f * (dl + dr), ca = .5f * (cl + cr);
  float pmid = .5f * (ql_ + qr_ + (ul - ur) * rhoa * ca);  // p_mid - P
  float c1 = (gamma + 1.f) / (2.f * gamma);
  float qL = (pmid <= ql_) ? 1.f : sqrtf(1.f + c1 * (pmid - ql_) * frcp(pl));
  float qR = (pmid <= qr_) ? 1.f : sqrtf(1.f + c1 * (pmid - qr_) * frcp(pr));
  float al = ul - cl * qL, ar = ur + cr * qR;
  float bp = ar > 0.f ? ar : 1e-20f;
  float bm = al < 0.f ? al : -1e-20f;
  float vxl = ul - al, vxr = ur - ar;
  float tl = ql_ + vxl * dl * ul;  // t_l - P
  float tr = qr_ + vxr * dr * ur;
  float ml = dl * vxl, mr = -(dr * vxr);
  float im = frcp(ml + mr);
  float am = (tl - tr) * im;
  float cp = (ml * tr + mr * tl) * im;  // contact pressure - P
  cp = cp > -P ? cp : -P;
  vxl = ul - bm;
  vxr = ur - bp;
  float fLd = dl * vxl, fRd = dr * vxr;
  float fLx = dl * ul * vxl + ql_, fRx = dr * ur * vxr + qr_;
  float fLy = dl * vl * vxl, fRy = dr * vr * vxr;
  float fLz = dl * zl * vxl, fRz = dr * zr * vxr;
  float fLe = el * vxl + pl * ul, fRe = er * vxr + pr * ur;
  float sl, sr, sm;
  if (am >= 0.f) {
    float i = frcp(am - bm);
    sl = am * i;
    sr = 0.f;
    sm = -bm * i;
  } else {
    float i = frcp(bp - am);
    sl = 0.f;
    sr = -am * i;
    sm = bp * i;
  }
  fl[IDN] = sl * fLd + sr * fRd;
  fl[ivx] = sl * fLx + sr * fRx + sm * cp;
  fl[ivy] = sl * fLy + sr * fRy;
  fl[ivz] = sl * fLz + sr * fRz;
  fl[IEN] = sl * fLe + sr * fRe + sm * (P + cp) * am;
}

// Shallow water about a base state b, the FP64 primitive state (h, -, v2, v3)
// of the face's left cell: the face states are b + dl and b + dr, with the
// deltas reconstructed from v - b (WENO5 is invariant under a constant shift),
// so in FP32 they carry only the local variation of the field.  The solvers
// return F(b + dl, b + dr) - F(b), every product involving b taken against a
// delta; F(b) = (h u, h u^2 + h^2/2, h u v) in the face frame is added back in
// FP64 by the update.  The large geopotential flux h^2/2 and its linear part
// h_b dh therefore never pass through FP32 as a whole.
EXO_HD void roe_sw_delta(int ivx, const float *b, const float *dl, const float *dr, float *fl) {
  int ivy = ivx == IVY ? IVZ : IVY;
  float hB = b[IDN], uB = b[ivx], vB = b[ivy];
  float hl = hB + dl[IDN], hr = hB + dr[IDN];
  float sl = sqrtf(hl), sr = sqrtf(hr);
  float isd = frcp(sl + sr);
  float ubar = ((uB + dl[ivx]) * sl + (uB + dr[ivx]) * sr) * isd;
  float vbar = ((vB + dl[ivy]) * sl + (vB + dr[ivy]) * sr) * isd;
  float cbar = sqrtf(0.5f * (hl + hr));
  float delh = dr[IDN] - dl[IDN];
  float delu = dr[ivx] - dl[ivx];
  float delv = dr[ivy] - dl[ivy];
  float hbar = sqrtf(hl * hr);
  float ic = frcp(cbar);
  float a1 = 0.5f * (cbar * delh - hbar * delu) * ic;
  float a2 = hbar * delv;
  float a3 = 0.5f * (cbar * delh + hbar * delu) * ic;
  float s0 = fabsf(ubar - cbar), s1 = fabsf(ubar), s2 = fabsf(ubar + cbar);
  // central part, 0.5 (F(b + dl) + F(b + dr)) - F(b)
  float fd = 0.f, fx = 0.f, fy = 0.f;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const float *d = side ? dr : dl;
    float dh = d[IDN], du = d[ivx], dv = d[ivy];
    float u2 = 2.f * uB + du;
    fd += hB * du + dh * (uB + du);
    fx += dh * uB * uB + hB * du * u2 + dh * du * u2 + hB * dh + 0.5f * dh * dh;
    fy += hB * (uB * dv + du * vB + du * dv) + dh * (uB + du) * (vB + dv);
  }
  fd *= 0.5f;
  fx *= 0.5f;
  fy *= 0.5f;
  fd -= 0.5f * (s0 * a1 + s2 * a3);
  fx -= 0.5f * (s0 * a1 * (ubar - cbar) + s2 * a3 * (ubar + cbar));
  fy -= 0.5f * (s0 * a1 * vbar + s1 * a2 + s2 * a3 * vbar);
  fl[IDN] = fd;
  fl[IVX] = 0.f;
  fl[ivx] = fx;
  fl[ivy] = fy;
}

EXO_HD void lmars_sw_delta(int ivx, const float *b, const float *dl, const float *dr, float *fl) {
  int ivy = ivx == IVY ? IVZ : IVY;
  float hB = b[IDN], uB = b[ivx], vB = b[ivy];
  float cbar = sqrtf(0.5f * ((hB + dl[IDN]) + (hB + dr[IDN])));
  float dhb = 0.5f * (dl[IDN] + dr[IDN]) + 0.5f * cbar * (dl[ivx] - dr[ivx]);          // hbar - hB
  float dub = 0.5f * (dl[ivx] + dr[ivx]) + 0.5f * frcp(cbar) * (dl[IDN] - dr[IDN]);  // ubar - uB
  const float *d = uB + dub > 0.f ? dl : dr;
  float dh = d[IDN], du = d[ivx], dv = d[ivy];
  float dm = uB * dh + dub * (hB + dh);  // ubar h - uB hB
  fl[IDN] = dm;
  fl[ivx] = dm * (uB + du) + uB * hB * du + hB * dhb + 0.5f * dhb * dhb;
  fl[ivy] = dm * (vB + dv) + uB * hB * dv;
  fl[IVX] = 0.f;
}

// Shallow-water face flux less the base flux, with the frame maps of the FP64
// path (the maps are linear, so the deltas map as the states do).  b, dl, dr
// are modified (face frame); mbase receives the base mass flux h_b u_b.
EXO_HD void riemann_sw_delta(int rs, int ivx, float *b, float *dl, float *dr, float *f, float cth,
                             float cosf, float sinf, float *mbase) {
  f[0] = f[1] = f[2] = f[3] = 0.f;
  local_frame<float>(ivx, cosf, sinf, b);
  local_frame<float>(ivx, cosf, sinf, dl);
  local_frame<float>(ivx, cosf, sinf, dr);
  if (rs == RS_ROE_SW) roe_sw_delta(ivx, b, dl, dr, f);
  else lmars_sw_delta(ivx, b, dl, dr, f);
  *mbase = b[IDN] * b[ivx];
  flux_to_global<float>(ivx, cosf, sinf, f);
  float ty = f[IVY], tz = f[IVZ];
  f[IVY] = ty + tz * cth;
  f[IVZ] = tz + ty * cth;
}

// Ideal-gas Riemann flux without the reference-pressure part, with the
// cubed-sphere frame maps of the FP64 path (PrimToLocal, FluxToGlobal,
// covariant factor).
EXO_HD void riemann_pert(int eos, int rs, int ivx, float gamma, float *wl, float *wr, float P,
                         float *f, float cth, float cosf, float sinf, float cosc) {
  f[0] = f[1] = f[2] = f[3] = 0.f;
  if (eos == EOS_IDEAL) f[4] = 0.f;
  if (ivx != IVX && rs != RS_LMARS) {
    local_frame<float>(ivx, cosf, sinf, wl);
    local_frame<float>(ivx, cosf, sinf, wr);
  }
  if (rs == RS_HLLC) hllc_pert(ivx, gamma, wl, wr, P, f, cosc);
  else lmars_pert(ivx, gamma, wl, wr, P, f);
  if (rs != RS_LMARS) {
    if (ivx != IVX) flux_to_global<float>(ivx, cosf, sinf, f);
    float ty = f[IVY], tz = f[IVZ];
    f[IVY] = ty + tz * cth;
    f[IVZ] = tz + ty * cth;
  }
}

}  // namespace exo
