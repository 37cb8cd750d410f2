// exo_setup.hpp -- host-side set-up of one cubed-sphere instance: mesh
// coordinates, metric tables, panel topology and the precomputed descriptors
// that the ghost-remap and edge-exchange kernels follow.  Everything here is
// evaluated once per instance in FP64 with the host libm, so the device never
// evaluates a trigonometric function.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/exocubed.h"

namespace exo {

// 2-D (per panel, nc3 x nc2) metric tables, in this order
enum Tab2 {
  T_COSC = 0,  // cosine_cell_kj_   gnomonic_equiangle.cpp:166
  T_SINC,      // sine_cell_kj_     :167
  T_COS2,      // cosine_face2_kj_  :169
  T_SIN2,      // sine_face2_kj_    :170
  T_COS3,      // cosine_face3_kj_  :172
  T_SIN3,      // sine_face3_kj_    :173
  T_DX2A,      // dx2f_ang_kj_      :182
  T_DX3A,      // dx3f_ang_kj_      :193
  T_DX2A_F3,   // dx2f_ang_face3_kj_:183
  T_DX3A_F2,   // dx3f_ang_face2_kj_:194
  T_CTH1,      // covariant factor of an x1 row (hllc_transform.cpp:238-251)
  T_CTH2,      // ... of an x2 face row (roe_shallow_yz.cpp:100-110)
  T_CTH3,      // ... of an x3 face row
  T_TANX,      // tan(x2v(j)), Coriolis and lat-lon maps
  T_TANY,      // tan(x3v(k))
  T_CORF,      // 2 Omega sin(lat) at the cell centre (W92.cpp:59)
  // global-frame momentum-flux factors of a unit normal pressure on an x2 /
  // x3 face (FluxToGlobal2/3 then the covariant projection; 1/0 for lmars.cpp,
  // which maps nothing): the mixed path adds P * factor in FP64
  T_PRY2, T_PRZ2, T_PRY3, T_PRZ3,
  T_NTAB
};

// 1-D (per panel, rows 0..nc2) tangent factors of the gnomonic coordinates:
// x = tan(x2v(j)), xf = tan(x2f(j)), y = tan(x3v(k)), yf = tan(x3f(k)) and
// C = sqrt(1 + x^2), CF, D = sqrt(1 + y^2), DF; the 2-D tables are formed
// from them (the cell-centre entries of the last row are unused)
enum Tab1 { T1_X = 0, T1_XF, T1_C, T1_CF, T1_Y, T1_YF, T1_D, T1_DF, T1_NTAB };

// One face-ghost cell of the panel-edge remap (cubed_sphere_utility.cpp:92-369
// as packed by PackData and unpacked by SetBoundarySameLevel).
struct GhostDesc {
  int tgt_panel, tgt_k, tgt_j;  // destination ghost cell (all i)
  int src_panel;                // source panel (0-based)
  int along_x2;                 // 1: source row is a k row, interpolated along j
  int src_row;                  // absolute source k (along_x2) or j
  int src_a0;                   // absolute source along index of the left point
  int src_block, tgt_block;     // block IDs for the vector maps
  int pad;
  double A, B, D;               // (y2-yq), (yq-y1), (y2-y1)
  double src_cy, src_cz, tgt_cy, tgt_cz;
  double sth_s, cth_s, sth_t, cth_t;
};

// One panel-edge face of the L/R exchange (cubed_sphere.cpp:326-687): the
// interior-side state of the source panel at its edge face, rotated into the
// neighbour's frame and stored in the neighbour's ghost-side slot.
struct EdgeDesc {
  int src_panel, src_dir;       // direction d: 0 ox2-, 1 ox2+, 2 ox3-, 3 ox3+
  int src_k, src_j;             // source face (face index on the normal axis)
  int tgt_panel, tgt_dir, tgt_a;
  int src_block, tgt_block;
  int pad;
  double x2s, x3s, x2d, x3d;    // tangents of the source / destination points
};

struct Setup {
  exo_config cfg;
  int N, nx1, nv, nt;
  int nc1, nc2, nc3;
  int is, ie, js, je, ks, ke;
  std::vector<double> x1f, x1v, dx1f;   // shared by all panels
  std::vector<double> x2f[6], x2v[6], x3f[6], x3v[6];
  std::vector<double> tab2;             // [T_NTAB][6][nc3][nc2]
  std::vector<double> tab1;             // [T1_NTAB][6][nc2 + 1]
  std::vector<GhostDesc> ghosts;        // 6 panels * 4 edges * 3 rows * N
  std::vector<EdgeDesc> edges;          // 6 panels * 4 edges * N

  size_t t1(int t, int p, int q) const { return ((size_t)t * 6 + p) * (nc2 + 1) + q; }
  size_t t2(int t, int p, int k, int j) const {
    return (((size_t)t * 6 + p) * nc3 + k) * nc2 + j;
  }
  double tab(int t, int p, int k, int j) const { return tab2[t2(t, p, k, j)]; }
  // geometry helpers identical to the device's
  double face1_area(int p, int k, int j, int i) const;
  double face2_area(int p, int k, int j, int i) const;
  double face3_area(int p, int k, int j, int i) const;
  double cell_volume(int p, int k, int j, int i) const;
  void latlon(int p, int k, int j, double *lat, double *lon) const;
};

// Builds everything above; returns an empty string or an error message.
std::string build_setup(const exo_config &cfg, Setup *s);

// Synthetic initial states of the benchmark configurations.
std::string init_state(const Setup &s, int kind, uint64_t seed, const double *params,
                       int nparams, double *u, double *sc);

}  // namespace exo
