// exo_setup.cpp -- host set-up of a cubed-sphere instance (one MeshBlock per
// panel, Athena mesh level 2): coordinates, metric tables, topology and the
// remap / edge-exchange descriptors consumed by the device kernels.
#include "exo_setup.hpp"

#include <cmath>
#include <cstring>

namespace exo {

static constexpr int NG = 3;
static constexpr double PI = 3.14159265358979323846;
enum { IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IEN = 4 };

// ---------------------------------------------------------------------------
// Athena's symmetric uniform mesh generator position of face `index` among
// `nrange` root cells, and the cubed-sphere panel maps that replace the mesh
// generator on x2 / x3 (src/exo3/cubed_sphere.cpp:37-73,
// patches/27.coordinates.patch:20-24,39-43).
static double mesh_fraction(long index, long nrange) {
  long noffset = index - nrange / 2;
  long noffset_ceil = index - (nrange + 1) / 2;
  return (double)(noffset + noffset_ceil) / (2.0 * nrange);
}

static double panel_x2(double x, int lx2) {
  double x_l = lx2 == 0 ? -0.5 : 0.0, x_u = lx2 == 0 ? 0.0 : 0.5;
  return (0.5 * (x - x_l) / (x_u - x_l) - 0.25) * PI;
}

static double panel_x3(double x, int lx3) {
  double x_l, x_u;
  switch (lx3) {
    case 0: x_l = -0.5; x_u = -1.0 / 6.0; break;
    case 1: x_l = -1.0 / 6.0; x_u = 1.0 / 6.0; break;
    default: x_l = 1.0 / 6.0; x_u = 0.5; break;
  }
  return (0.5 * (x - x_l) / (x_u - x_l) - 0.25) * PI;
}

// RLLFromXYP, src/exo3/cubed_sphere_transform.cpp:63-118
static void rll_from_xyp(double dX, double dY, int nP, double *lon, double *lat) {
  switch (nP) {
    case 1: *lon = atan(dX); *lat = atan(dY / sqrt(1.0 + dX * dX)); break;
    case 3: *lon = atan(dX) + 0.5 * PI; *lat = atan(dY / sqrt(1.0 + dX * dX)); break;
    case 5: *lon = atan(dX) + PI; *lat = atan(dY / sqrt(1.0 + dX * dX)); break;
    case 2: *lon = atan(dX) + 1.5 * PI; *lat = atan(dY / sqrt(1.0 + dX * dX)); break;
    case 0:
      if (fabs(dX) > 1.0e-10) *lon = atan2(dX, -dY);
      else if (dY <= 0.0) *lon = 0.0;
      else *lon = PI;
      *lat = 0.5 * PI - atan(sqrt(dX * dX + dY * dY));
      break;
    default:
      if (fabs(dX) > 1.0e-10) *lon = atan2(dX, dY);
      else if (dY > 0.0) *lon = 0.0;
      else *lon = PI;
      *lat = -0.5 * PI + atan(sqrt(dX * dX + dY * dY));
      break;
  }
  if (*lon < 0.0) *lon += 2.0 * PI;
}

// host copies of the panel / lat-lon vector maps (cubed_sphere_transform.cpp:11-61)
static void abp_from_rll(double X, double Y, int b, double U, double V, double *V2,
                         double *V3) {
  double C = sqrt(1 + X * X), D = sqrt(1 + Y * Y), E = sqrt(X * X + Y * Y);
  double delta = sqrt(1 + X * X + Y * Y);
  if (b == 2 || b == 3 || b == 4 || b == 6) {
    *V2 = X * Y / delta * U - V;
    *V3 = C * D / delta * U;
  } else if (b == 1) {
    *V2 = -D * Y / (delta * E) * U - D * X / E * V;
    *V3 = C * X / (delta * E) * U - C * Y / E * V;
  } else {
    *V2 = D * Y / (delta * E) * U + D * X / E * V;
    *V3 = -C * X / (delta * E) * U + C * Y / E * V;
  }
}

double Setup::face1_area(int p, int k, int j, int i) const {
  return tab(T_DX2A, p, k, j) * tab(T_DX3A, p, k, j) * x1f[i] * x1f[i] * tab(T_SINC, p, k, j);
}
double Setup::face2_area(int p, int k, int j, int i) const {
  return x1v[i] * tab(T_DX3A_F2, p, k, j) * dx1f[i];
}
double Setup::face3_area(int p, int k, int j, int i) const {
  return x1v[i] * tab(T_DX2A_F3, p, k, j) * dx1f[i];
}
double Setup::cell_volume(int p, int k, int j, int i) const {
  return dx1f[i] * tab(T_DX2A, p, k, j) * tab(T_DX3A, p, k, j) * x1v[i] * x1v[i] *
         tab(T_SINC, p, k, j);
}

// CubedSphere::GetLatLon, cubed_sphere.cpp:78-90
void Setup::latlon(int p, int k, int j, double *lat, double *lon) const {
  double dX = tan(x2v[p][j]);
  double dY = tan(x3v[p][k]);
  rll_from_xyp(dY, -dX, p, lon, lat);
  *lon = 2.0 * PI - *lon;
}

// ---------------------------------------------------------------------------
// topology tables: src/exo3/cubed_sphere_utility.cpp:378-391 (dinv, tgside,
// tgbid), the PackDataR* choice per (block, side) of :410-614, and
// CubedSphere::TransformOX (cs_transform_ox.cpp:11-229) at mesh level 2.
// Side index d: 0 = ox2 -1, 1 = ox2 +1, 2 = ox3 -1, 3 = ox3 +1.
static const int DINV[6][4] = {{1, 0, 0, 1}, {0, 0, 0, 0}, {0, 1, 0, 0},
                               {1, 0, 0, 0}, {0, 1, 1, 0}, {1, 1, 0, 0}};
static const int TGSIDE[6][4] = {{-1, -1, -1, -1}, {1, -1, 1, -1}, {-1, -1, 1, -1},
                                 {1, 1, 1, -1},    {1, 1, 1, 1},   {-1, 1, 1, -1}};
static const int TGBID[6][4] = {{6, 2, 3, 4}, {1, 5, 3, 4}, {1, 5, 6, 2},
                                {1, 5, 2, 6}, {2, 6, 3, 4}, {1, 5, 4, 3}};
static const int PACKR[6][4] = {{2, 0, 3, 1}, {0, 0, 0, 0}, {1, 3, 0, 0},
                                {3, 1, 0, 0}, {0, 2, 1, 3}, {2, 2, 0, 0}};
// neighbour across side d of block b and the side of the neighbour that
// faces back (as a side index of the neighbour)
static const int NBR_BLOCK[6][4] = {{6, 2, 3, 4}, {1, 5, 3, 4}, {1, 5, 6, 2},
                                    {1, 5, 2, 6}, {2, 6, 3, 4}, {1, 5, 4, 3}};
static const int NBR_SIDE[6][4] = {{0, 0, 0, 0}, {1, 0, 3, 2}, {2, 2, 3, 2},
                                   {3, 3, 3, 2}, {1, 1, 1, 1}, {0, 1, 3, 2}};

static int side_ox2(int d) { return d == 0 ? -1 : (d == 1 ? 1 : 0); }
static int side_ox3(int d) { return d == 2 ? -1 : (d == 3 ? 1 : 0); }

// CalculateInterpLocations, cubed_sphere_utility.cpp:70-90
static double interp_loc(int loc_n, int N_blk, int k, bool ghost) {
  double temp0 = tan((double)1.0 * (N_blk - 1 - 2 * k) / (N_blk * 4) * PI);
  if (ghost) {
    temp0 = 1.0 / temp0;
    temp0 = 1.0 + temp0 * temp0;
  } else {
    temp0 = 1.0 + temp0 * temp0;
  }
  double t = pow(tan((double)1.0 * (2 * loc_n + 1) / (N_blk * 4) * PI), 2.0);
  if (loc_n < 0) return -acos(sqrt(temp0 / (temp0 + t)));
  return acos(sqrt(temp0 / (temp0 + t)));
}

// positions (kk, jj) of a (nk x nj) array in PackDataR<r> order
static void pack_order(int r, int nk, int nj, std::vector<std::pair<int, int>> *out) {
  out->clear();
  if (r == 0) {
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j) out->push_back({k, j});
  } else if (r == 1) {
    for (int j = nj - 1; j >= 0; --j)
      for (int k = 0; k < nk; ++k) out->push_back({k, j});
  } else if (r == 2) {
    for (int k = nk - 1; k >= 0; --k)
      for (int j = nj - 1; j >= 0; --j) out->push_back({k, j});
  } else {
    for (int j = 0; j < nj; ++j)
      for (int k = nk - 1; k >= 0; --k) out->push_back({k, j});
  }
}

static void build_ghosts(Setup *s) {
  int N = s->N;
  std::vector<double> src_x(N), src_coord(N), tgt_x(N), tgt_coord(N);
  std::vector<std::pair<int, int>> order;
  for (int p = 0; p < 6; ++p) {
    int blockID = p + 1;
    for (int d = 0; d < 4; ++d) {
      int ox2 = side_ox2(d), ox3 = side_ox3(d);
      int sj = ox2 > 0 ? s->je - NG + 1 : s->js;
      int ej = ox2 < 0 ? s->js + NG - 1 : s->je;
      int sk = ox3 > 0 ? s->ke - NG + 1 : s->ks;
      int ek = ox3 < 0 ? s->ks + NG - 1 : s->ke;
      bool along_x2 = ox3 != 0;
      int nk = ek - sk + 1, nj = ej - sj + 1;
      int DirInv = DINV[p][d], TgtSide = TGSIDE[p][d], TgtID = TGBID[p][d];
      // interpolation descriptors of the (nk x nj) interpolated array
      std::vector<GhostDesc> itp((size_t)nk * nj);
      for (int row = 0; row < 3; ++row) {
        int k_now, SrcSide;
        if (along_x2) {
          int k = sk + row;
          if (sk < ej - NG) { k_now = k - sk; SrcSide = -1; }
          else { k_now = ek - k; SrcSide = 1; }
        } else {
          int j = sj + row;
          if (sj < ek - NG) { k_now = j - sj; SrcSide = -1; }
          else { k_now = ej - j; SrcSide = 1; }
        }
        int n_start = -N / 2;
        for (int a = 0; a < N; ++a) {
          src_x[a] = interp_loc(a - N / 2, N, k_now, false);
          src_coord[a] = PI / 2.0 / N * (a + 0.5) - PI / 4.0;
        }
        for (int a = 0; a < N; ++a) {
          tgt_x[a] = interp_loc(n_start + a, N, k_now, true);
          tgt_coord[a] = DirInv == 1 ? -src_coord[n_start + N / 2 + a]
                                     : src_coord[n_start + N / 2 + a];
        }
        int sp = 0;
        for (int a = 0; a < N; ++a) {
          while (sp + 2 < N && tgt_x[a] > src_x[sp + 1]) sp++;
          double y1 = src_x[sp], y2 = src_x[sp + 1], yq = tgt_x[a];
          GhostDesc g;
          std::memset(&g, 0, sizeof(g));
          g.src_panel = p;
          g.along_x2 = along_x2 ? 1 : 0;
          g.src_row = along_x2 ? sk + row : sj + row;
          g.src_a0 = sp + (along_x2 ? sj : sk);
          g.src_block = blockID;
          g.tgt_block = TgtID;
          g.A = y2 - yq;
          g.B = yq - y1;
          g.D = y2 - y1;
          double tgt_cy, tgt_cz, src_cy, src_cz;
          if (along_x2) {
            tgt_cy = tan(tgt_coord[a]);
            tgt_cz = tan(TgtSide * (PI / 2.0 / (N) * (k_now + 0.5) + PI / 4.0));
            src_cz = tan(SrcSide * (-PI / 2.0 / (N) * (k_now + 0.5) + PI / 4.0));
            src_cy = tan(((y2 - yq) * src_coord[sp] + (yq - y1) * src_coord[sp + 1]) /
                         (y2 - y1));
          } else {
            tgt_cz = tan(tgt_coord[a]);
            tgt_cy = tan(TgtSide * (PI / 2.0 / (N) * (k_now + 0.5) + PI / 4.0));
            src_cy = tan(SrcSide * (-PI / 2.0 / (N) * (k_now + 0.5) + PI / 4.0));
            src_cz = tan(((y2 - yq) * src_coord[sp] + (yq - y1) * src_coord[sp + 1]) /
                         (y2 - y1));
            // InteprolateX3's exchange of the target tangents for these pairs
            // (cubed_sphere_utility.cpp:318-329)
            if (((TgtID == 1) && (blockID == 3)) || ((TgtID == 1) && (blockID == 4)) ||
                ((TgtID == 5) && (blockID == 3)) || ((TgtID == 5) && (blockID == 4)) ||
                ((TgtID == 3) && (blockID == 1)) || ((TgtID == 4) && (blockID == 1)) ||
                ((TgtID == 3) && (blockID == 5)) || ((TgtID == 4) && (blockID == 5))) {
              double t = tgt_cy;
              tgt_cy = tgt_cz;
              tgt_cz = t;
            }
          }
          g.src_cy = src_cy;
          g.src_cz = src_cz;
          g.tgt_cy = tgt_cy;
          g.tgt_cz = tgt_cz;
          g.sth_s = sqrt(1 + src_cy * src_cy + src_cz * src_cz) / sqrt(1 + src_cy * src_cy) /
                    sqrt(1 + src_cz * src_cz);
          g.cth_s = -src_cy * src_cz / sqrt(1 + src_cy * src_cy) / sqrt(1 + src_cz * src_cz);
          g.sth_t = sqrt(1 + tgt_cy * tgt_cy + tgt_cz * tgt_cz) / sqrt(1 + tgt_cy * tgt_cy) /
                    sqrt(1 + tgt_cz * tgt_cz);
          g.cth_t = -tgt_cy * tgt_cz / sqrt(1 + tgt_cy * tgt_cy) / sqrt(1 + tgt_cz * tgt_cz);
          size_t pos = along_x2 ? (size_t)row * nj + a : (size_t)a * nj + row;
          itp[pos] = g;
        }
      }
      // pack (PackDataR*) -> buffer -> unpack into the neighbour's ghost zone
      pack_order(PACKR[p][d], nk, nj, &order);
      int tb = NBR_BLOCK[p][d] - 1, td = NBR_SIDE[p][d];
      int tox2 = side_ox2(td), tox3 = side_ox3(td);
      int tsj, tej, tsk, tek;
      if (tox2 == 0) { tsj = s->js; tej = s->je; }
      else if (tox2 > 0) { tsj = s->je + 1; tej = s->je + NG; }
      else { tsj = s->js - NG; tej = s->js - 1; }
      if (tox3 == 0) { tsk = s->ks; tek = s->ke; }
      else if (tox3 > 0) { tsk = s->ke + 1; tek = s->ke + NG; }
      else { tsk = s->ks - NG; tek = s->ks - 1; }
      size_t q = 0;
      for (int k = tsk; k <= tek; ++k)
        for (int j = tsj; j <= tej; ++j) {
          GhostDesc g = itp[(size_t)order[q].first * nj + order[q].second];
          g.tgt_panel = tb;
          g.tgt_k = k;
          g.tgt_j = j;
          s->ghosts.push_back(g);
          ++q;
        }
    }
  }
}

static void build_edges(Setup *s) {
  int N = s->N;
  for (int p = 0; p < 6; ++p) {
    int blockID = p + 1;
    for (int d = 0; d < 4; ++d) {
      int ox2 = side_ox2(d), ox3 = side_ox3(d);
      int tb = NBR_BLOCK[p][d] - 1, td = NBR_SIDE[p][d];
      int tox2 = side_ox2(td), tox3 = side_ox3(td);
      int inv = DINV[p][d];
      // sender's traversal of its edge faces (cubed_sphere.cpp:419-554)
      std::vector<std::pair<int, int>> src;
      for (int q = 0; q < N; ++q) {
        int a = inv ? N - 1 - q : q;
        if (ox2 == 1) src.push_back({s->ks + a, s->je + 1});
        else if (ox2 == -1) src.push_back({s->ks + a, s->js});
        else if (ox3 == 1) src.push_back({s->ke + 1, s->js + a});
        else src.push_back({s->ks, s->js + a});
      }
      for (int q = 0; q < N; ++q) {
        int k = src[q].first, j = src[q].second;
        EdgeDesc e;
        std::memset(&e, 0, sizeof(e));
        e.src_panel = p;
        e.src_dir = d;
        e.src_k = k;
        e.src_j = j;
        e.tgt_panel = tb;
        e.tgt_dir = td;
        e.tgt_a = q;  // receiver unpacks its edge in ascending order
        e.src_block = blockID;
        e.tgt_block = TGBID[p][d];
        double x2s, x3s, x2d, x3d, pos_along;
        if (ox2 == 1) { x2s = PI / 4.0; x3s = s->x3v[p][k]; pos_along = x3s; }
        else if (ox2 == -1) { x2s = -PI / 4.0; x3s = s->x3v[p][k]; pos_along = x3s; }
        else if (ox3 == 1) { x2s = s->x2v[p][j]; x3s = PI / 4.0; pos_along = x2s; }
        else { x2s = s->x2v[p][j]; x3s = -PI / 4.0; pos_along = x2s; }
        double pa = inv ? -pos_along : pos_along;
        if (tox2 == 1) { x2d = PI / 4.0; x3d = pa; }
        else if (tox2 == -1) { x2d = -PI / 4.0; x3d = pa; }
        else if (tox3 == 1) { x2d = pa; x3d = PI / 4.0; }
        else { x2d = pa; x3d = -PI / 4.0; }
        e.x2d = tan(x2d);
        e.x3d = tan(x3d);
        e.x2s = tan(x2s);
        e.x3s = tan(x3s);
        s->edges.push_back(e);
      }
    }
  }
}

std::string build_setup(const exo_config &cfg, Setup *s) {
  if (cfg.nx < 4 || (cfg.nx % 2) != 0) return "nx must be an even number >= 4";
  if (cfg.nx1 < 1) return "nx1 must be >= 1";
  if (cfg.eos != EXO_EOS_SHALLOW_YZ && cfg.eos != EXO_EOS_IDEAL) return "unknown eos";
  if (cfg.rsolver < EXO_RS_ROE_SW || cfg.rsolver > EXO_RS_LMARS) return "unknown rsolver";
  if (cfg.eos == EXO_EOS_SHALLOW_YZ && cfg.rsolver > EXO_RS_LMARS_SW)
    return "shallow-water EOS needs a shallow-water Riemann solver";
  if (cfg.eos == EXO_EOS_IDEAL && cfg.rsolver < EXO_RS_HLLC)
    return "ideal-gas EOS needs the HLLC or LMARS Riemann solver";
  // bit 1: x1 column solve; bit 8: the full 5x5 system instead of the partial
  // 3x3 one (solve_implicit_3d.cpp:102-135).  The x2/x3 bits are not supported.
  if (cfg.implicit_flag & ~9) return "implicit_flag must be 0, 1, 8 or 9 (x1 column correction)";
  if (cfg.eos == EXO_EOS_SHALLOW_YZ && cfg.implicit_flag != 0)
    return "the implicit solve needs the ideal-gas EOS";
  if (cfg.eos == EXO_EOS_IDEAL && cfg.nx1 < 3) return "ideal gas needs nx1 >= 3";
  if (cfg.ntracer < 0 || cfg.ntracer > 4) return "ntracer must be 0..4";
  if (!(cfg.x1max > cfg.x1min) || !(cfg.x1min > 0)) return "bad radial range";
  if (!(cfg.cfl > 0)) return "cfl must be positive";
  if (cfg.precision != EXO_PREC_FP64 && cfg.precision != EXO_PREC_MIXED)
    return "unknown precision";
  if (cfg.coriolis < EXO_COR_NONE || cfg.coriolis > EXO_COR_3D) return "unknown coriolis";
  s->cfg = cfg;
  s->N = cfg.nx;
  s->nx1 = cfg.nx1;
  s->nv = cfg.eos == EXO_EOS_SHALLOW_YZ ? 4 : 5;
  s->nt = cfg.ntracer;
  s->nc1 = cfg.nx1 + 2 * NG;
  s->nc2 = s->nc3 = cfg.nx + 2 * NG;
  s->is = s->js = s->ks = NG;
  s->ie = NG + cfg.nx1 - 1;
  s->je = s->ke = NG + cfg.nx - 1;
  int N = s->N, nx1 = s->nx1, nc1 = s->nc1, nc2 = s->nc2, nc3 = s->nc3;
  // x1: uniform symmetric generator, block edges pinned (Athena Coordinates)
  s->x1f.assign(nc1 + 1, 0.);
  s->x1v.assign(nc1, 0.);
  s->dx1f.assign(nc1, 0.);
  for (int i = 0; i <= nc1; ++i) {
    double rx = mesh_fraction(i - NG, nx1);
    s->x1f[i] = (0.5 - rx) * cfg.x1min + (0.5 + rx) * cfg.x1max;
  }
  s->x1f[NG] = cfg.x1min;
  s->x1f[NG + nx1] = cfg.x1max;
  double dx1 = (cfg.x1max - cfg.x1min) / nx1;
  for (int i = 0; i < nc1; ++i) s->dx1f[i] = dx1;
  for (int i = 0; i < nc1; ++i) s->x1v[i] = 0.5 * (s->x1f[i + 1] + s->x1f[i]);
  s->tab2.assign((size_t)T_NTAB * 6 * nc3 * nc2, 0.);
  s->tab1.assign((size_t)T1_NTAB * 6 * (nc2 + 1), 0.);
  for (int p = 0; p < 6; ++p) {
    int lx2 = p % 2, lx3 = p / 2;
    auto &x2f = s->x2f[p], &x2v = s->x2v[p], &x3f = s->x3f[p], &x3v = s->x3v[p];
    x2f.assign(nc2 + 1, 0.);
    x2v.assign(nc2, 0.);
    x3f.assign(nc3 + 1, 0.);
    x3v.assign(nc3, 0.);
    for (int j = 0; j <= nc2; ++j)
      x2f[j] = panel_x2(mesh_fraction((long)(j - NG) + (long)lx2 * N, 2L * N), lx2);
    for (int j = 0; j < nc2; ++j) x2v[j] = 0.5 * (x2f[j + 1] + x2f[j]);
    for (int k = 0; k <= nc3; ++k)
      x3f[k] = panel_x3(mesh_fraction((long)(k - NG) + (long)lx3 * N, 3L * N), lx3);
    for (int k = 0; k < nc3; ++k) x3v[k] = 0.5 * (x3f[k + 1] + x3f[k]);
    double *T1 = s->tab1.data();
    for (int q = 0; q <= nc2; ++q) {
      double xf = tan(x2f[q]), yf = tan(x3f[q]);
      T1[s->t1(T1_XF, p, q)] = xf;
      T1[s->t1(T1_CF, p, q)] = sqrt(1.0 + xf * xf);
      T1[s->t1(T1_YF, p, q)] = yf;
      T1[s->t1(T1_DF, p, q)] = sqrt(1.0 + yf * yf);
      if (q == nc2) break;
      double x = tan(x2v[q]), y = tan(x3v[q]);
      T1[s->t1(T1_X, p, q)] = x;
      T1[s->t1(T1_C, p, q)] = sqrt(1.0 + x * x);
      T1[s->t1(T1_Y, p, q)] = y;
      T1[s->t1(T1_D, p, q)] = sqrt(1.0 + y * y);
    }
    // GnomonicEquiangle constructor, gnomonic_equiangle.cpp:154-196
    for (int k = 0; k < nc3; ++k)
      for (int j = 0; j < nc2; ++j) {
        double x = T1[s->t1(T1_X, p, j)];
        double xf = T1[s->t1(T1_XF, p, j)];
        double y = T1[s->t1(T1_Y, p, k)];
        double yf = T1[s->t1(T1_YF, p, k)];
        double C = T1[s->t1(T1_C, p, j)];
        double Cf = T1[s->t1(T1_CF, p, j)];
        double D = T1[s->t1(T1_D, p, k)];
        double Df = T1[s->t1(T1_DF, p, k)];
        double *T = s->tab2.data();
        T[s->t2(T_COSC, p, k, j)] = -x * y / (C * D);
        T[s->t2(T_SINC, p, k, j)] = sqrt(1.0 + x * x + y * y) / (C * D);
        T[s->t2(T_COS2, p, k, j)] = -xf * y / (Cf * D);
        T[s->t2(T_SIN2, p, k, j)] = sqrt(1.0 + xf * xf + y * y) / (Cf * D);
        T[s->t2(T_COS3, p, k, j)] = -x * yf / (C * Df);
        T[s->t2(T_SIN3, p, k, j)] = sqrt(1.0 + x * x + yf * yf) / (C * Df);
        double x1 = tan(x2f[j]);
        double x2 = tan(x2f[j + 1]);
        double delta1 = sqrt(1.0 + x1 * x1 + y * y);
        double delta2 = sqrt(1.0 + x2 * x2 + y * y);
        double delta1f = sqrt(1.0 + x1 * x1 + yf * yf);
        double delta2f = sqrt(1.0 + x2 * x2 + yf * yf);
        T[s->t2(T_DX2A, p, k, j)] = acos((1 + x1 * x2 + y * y) / (delta1 * delta2));
        T[s->t2(T_DX2A_F3, p, k, j)] = acos((1 + x1 * x2 + yf * yf) / (delta1f * delta2f));
        double y1 = tan(x3f[k]);
        double y2 = tan(x3f[k + 1]);
        delta1 = sqrt(1.0 + x * x + y1 * y1);
        delta2 = sqrt(1.0 + x * x + y2 * y2);
        delta1f = sqrt(1.0 + xf * xf + y1 * y1);
        delta2f = sqrt(1.0 + xf * xf + y2 * y2);
        T[s->t2(T_DX3A, p, k, j)] = acos((1 + x * x + y1 * y2) / (delta1 * delta2));
        T[s->t2(T_DX3A_F2, p, k, j)] = acos((1 + xf * xf + y1 * y2) / (delta1f * delta2f));
        // solvers' covariant projection factor: cth = -x * y / C / D with the
        // row's (x, y) (e.g. roe_shallow_yz.cpp:100-110, hllc_transform.cpp:238-251)
        {
          double xx = x, yy = y;
          T[s->t2(T_CTH1, p, k, j)] = -xx * yy / sqrt(1. + xx * xx) / sqrt(1. + yy * yy);
          xx = xf;
          yy = y;
          T[s->t2(T_CTH2, p, k, j)] = -xx * yy / sqrt(1. + xx * xx) / sqrt(1. + yy * yy);
          xx = x;
          yy = yf;
          T[s->t2(T_CTH3, p, k, j)] = -xx * yy / sqrt(1. + xx * xx) / sqrt(1. + yy * yy);
        }
        T[s->t2(T_TANX, p, k, j)] = x;
        T[s->t2(T_TANY, p, k, j)] = y;
        // pressure-flux factors: local normal flux P -> global covariant flux
        if (cfg.rsolver == EXO_RS_LMARS) {
          T[s->t2(T_PRY2, p, k, j)] = 1.;
          T[s->t2(T_PRZ2, p, k, j)] = 0.;
          T[s->t2(T_PRY3, p, k, j)] = 0.;
          T[s->t2(T_PRZ3, p, k, j)] = 1.;
        } else {
          double s2 = T[s->t2(T_SIN2, p, k, j)], c2 = T[s->t2(T_COS2, p, k, j)];
          double s3 = T[s->t2(T_SIN3, p, k, j)], c3 = T[s->t2(T_COS3, p, k, j)];
          double gi2 = 1.0 / (s2 * s2), gi3 = 1.0 / (s3 * s3);
          double T22 = sqrt(gi2), T32 = -sqrt(gi2) * c2;  // FluxToGlobal2
          double T23 = -c3 * sqrt(gi3), T33 = sqrt(gi3);  // FluxToGlobal3
          double ch2 = T[s->t2(T_CTH2, p, k, j)], ch3 = T[s->t2(T_CTH3, p, k, j)];
          T[s->t2(T_PRY2, p, k, j)] = T22 + T32 * ch2;
          T[s->t2(T_PRZ2, p, k, j)] = T32 + T22 * ch2;
          T[s->t2(T_PRY3, p, k, j)] = T23 + T33 * ch3;
          T[s->t2(T_PRZ3, p, k, j)] = T33 + T23 * ch3;
        }
      }
  }
  for (int p = 0; p < 6; ++p)
    for (int k = 0; k < nc3; ++k)
      for (int j = 0; j < nc2; ++j) {
        double lat, lon;
        s->latlon(p, k, j, &lat, &lon);
        s->tab2[s->t2(T_CORF, p, k, j)] = 2. * cfg.omega * sin(lat);
      }
  s->ghosts.clear();
  s->edges.clear();
  build_ghosts(s);
  build_edges(s);
  return "";
}

// ---------------------------------------------------------------------------
// Synthetic initial states.  SplitMix64 stream -> uniform [0,1) -> Box-Muller.
namespace {
struct Rng {
  uint64_t st;
  uint64_t next() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
  }
};
}  // namespace

std::string init_state(const Setup &s, int kind, uint64_t seed, const double *prm,
                       int nprm, double *u, double *sc) {
  int N = s.N, nc1 = s.nc1, nc2 = s.nc2, nc3 = s.nc3, nv = s.nv;
  auto U4 = [&](int p, int n, int k, int j, int i) -> double & {
    return u[((((size_t)p * nv + n) * nc3 + k) * nc2 + j) * nc1 + i];
  };
  std::memset(u, 0, sizeof(double) * 6 * nv * nc3 * nc2 * nc1);
  Rng rng{seed};
  auto P = [&](int i, double dflt) { return i < nprm ? prm[i] : dflt; };
  if (kind == 0) {
    if (s.cfg.eos != EXO_EOS_SHALLOW_YZ) return "kind 0 needs the shallow-water EOS";
    double h0 = P(0, 5960.), amp = P(1, 50.), g = P(2, 9.81);
    int nmodes = (int)P(3, 16.);
    double eps = sin(P(4, 5.0) * PI / 180.);
    double a = s.x1v[s.is], om = s.cfg.omega;
    std::vector<int> km(nmodes), lm(nmodes);
    std::vector<double> al(nmodes);
    for (int m = 0; m < nmodes; ++m) {
      km[m] = (int)(rng.next() % 5);
      lm[m] = (int)(rng.next() % 5);
      al[m] = 2.0 * PI * rng.uniform();
    }
    for (int p = 0; p < 6; ++p)
      for (int k = s.ks; k <= s.ke; ++k)
        for (int j = s.js; j <= s.je; ++j) {
          double lat, lon;
          s.latlon(p, k, j, &lat, &lon);
          double C = cos(lat), z = sin(lat), th = acos(z);
          double h = h0, dhdphi = 0., dhdlam_over_c = 0.;
          for (int m = 0; m < nmodes; ++m) {
            int kk = km[m], ll = lm[m];
            double ck = pow(C, kk), ck1 = kk > 0 ? pow(C, kk - 1) : 0.;
            double cz = cos(kk * lon + al[m]), sz = sin(kk * lon + al[m]);
            double T = cos(ll * th), dT = ll * sin(ll * th);  // dT/dphi = T'(z) cos(lat)
            h += amp * ck * cz * T;
            dhdphi += amp * (-(double)kk * ck1 * z * cz * T + ck * cz * dT);
            dhdlam_over_c += amp * (-(double)kk * ck1 * sz * T);
          }
          // geostrophic wind, smoothly switched off across the equator
          double gf = g * z / (2. * om * (z * z + eps * eps));
          double Uw = -gf / a * dhdphi, Vw = gf / a * dhdlam_over_c;
          double X = tan(s.x2v[p][j]), Y = tan(s.x3v[p][k]);
          double v2, v3;
          abp_from_rll(X, Y, p + 1, -Uw, Vw, &v2, &v3);  // CubedSphere::GetVyVz
          double cth = s.tab(T_COSC, p, k, j);
          double phi = g * h;
          for (int i = s.is; i <= s.ie; ++i) {
            U4(p, IDN, k, j, i) = phi;
            U4(p, IVX, k, j, i) = 0.;
            // shallow_yz_hydro.cpp:76-98 PrimitiveToConserved
            U4(p, IVY, k, j, i) = (v2 + v3 * cth) * phi;
            U4(p, IVZ, k, j, i) = (v3 + v2 * cth) * phi;
          }
        }
    (void)N;
    return "";
  }
  if (s.cfg.eos != EXO_EOS_IDEAL) return "kinds 1 and 2 need the ideal-gas EOS";
  double g = -s.cfg.grav, gam = s.cfg.gamma;
  double Rd, ps, noise;
  std::vector<double> Tz(nc1), pz(nc1);
  double wave_amp = 0.;
  bool noise_is_kelvin;
  int qpar;
  if (kind == 1) {
    double T0 = P(0, 250.);
    ps = P(1, 1e5);
    Rd = P(2, 287.);
    noise = P(3, 0.01);
    wave_amp = P(4, 5.);
    noise_is_kelvin = false;
    qpar = 5;
    for (int i = 0; i < nc1; ++i) {
      double z = s.x1v[i] - s.cfg.x1min;
      Tz[i] = T0;
      pz[i] = ps * exp(-g * z / (Rd * T0));
    }
  } else if (kind == 2) {
    double Ts = P(0, 300.), lapse = P(1, 6.5e-3), Tst = P(2, 200.);
    ps = P(3, 1e5);
    Rd = P(4, 287.);
    noise = P(5, 0.5);
    noise_is_kelvin = true;
    qpar = 6;
    double zt = (Ts - Tst) / lapse;
    double pt = ps * pow(Tst / Ts, g / (Rd * lapse));
    for (int i = 0; i < nc1; ++i) {
      double z = s.x1v[i] - s.cfg.x1min;
      if (z < zt) {
        Tz[i] = Ts - lapse * z;
        pz[i] = ps * pow(Tz[i] / Ts, g / (Rd * lapse));
      } else {
        Tz[i] = Tst;
        pz[i] = pt * exp(-g * (z - zt) / (Rd * Tst));
      }
    }
  } else {
    return "unknown initial-state kind";
  }
  double ph[4];
  for (int m = 0; m < 4; ++m) ph[m] = 2.0 * PI * rng.uniform();
  double qv_mean = P(qpar, 3e-3), qv_sig = P(qpar + 1, 0.5);
  for (int p = 0; p < 6; ++p)
    for (int k = s.ks; k <= s.ke; ++k)
      for (int j = s.js; j <= s.je; ++j) {
        double lat, lon;
        s.latlon(p, k, j, &lat, &lon);
        double Uw = 0.;
        for (int m = 0; m < 4; ++m) Uw += 0.25 * wave_amp * cos((m + 1) * lon + ph[m]);
        Uw *= cos(lat);
        double X = tan(s.x2v[p][j]), Y = tan(s.x3v[p][k]);
        double v2, v3;
        abp_from_rll(X, Y, p + 1, -Uw, 0., &v2, &v3);
        double cth = s.tab(T_COSC, p, k, j);
        for (int i = s.is; i <= s.ie; ++i) {
          double T = Tz[i] + (noise_is_kelvin ? noise : noise * Tz[i]) * rng.normal();
          double pr = pz[i], rho = pr / (Rd * T);
          double m1 = 0. * rho, m2 = v2 * rho, m3 = v3 * rho;
          // ideal_moist_hydro.cpp:168-223 PrimitiveToConserved (NVAPOR = 0)
          double c2 = m2 + m3 * cth, c3 = m3 + m2 * cth;
          double KE = 0.5 * (m1 * 0. + c2 * v2 + c3 * v3);
          U4(p, IDN, k, j, i) = rho;
          U4(p, IVX, k, j, i) = m1;
          U4(p, IVY, k, j, i) = c2;
          U4(p, IVZ, k, j, i) = c3;
          U4(p, IEN, k, j, i) = 1.0 / (gam - 1.0) * pr * 1. / 1. + KE;
          if (sc) {
            for (int t = 0; t < s.nt; ++t) {
              double r = 0.;
              if (t == 0) {
                double mu = log(qv_mean) - 0.5 * qv_sig * qv_sig;
                r = exp(mu + qv_sig * rng.normal());
              }
              sc[((((size_t)p * s.nt + t) * nc3 + k) * nc2 + j) * nc1 + i] = rho * r;
            }
          }
        }
      }
  return "";
}

}  // namespace exo
