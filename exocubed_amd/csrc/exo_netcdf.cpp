// exo_netcdf.cpp -- per-MeshBlock NetCDF output in the reference's layout.
//
// What is written follows NetcdfOutput::WriteOutputFile
// (src/outputs/netcdf.cpp:34-469):
//   * one file per MeshBlock, <basename>.block<gid>.<file_id>.<number %05d>.nc
//     (:73-90);
//   * dimensions time (record), x1, x1f, x2, x2f, x3, x3f, the face dimensions
//     only where a direction has more than one cell (:100-121);
//   * float coordinate variables with their axis / units / long_name /
//     domain_decomposition attributes (:124-203), the global NumFilesInSet
//     (:205);
//   * every output field as float (time, x1, x2, x3), x3 fastest (:227-279 and
//     the cell-centred branch :439-447), units and long_name from the metadata
//     table of src/outputs/output_utils.cpp:29-48.
//
// The reference links libnetcdf and asks for NC_NETCDF4 (HDF5) files (:94).
// This writer emits the NetCDF-3 64-bit-offset format instead (the classic
// format's published specification, big-endian), which the netCDF library and
// its Python bindings read through the same API.  One deliberate difference:
// the reference fills the x3f coordinate from x1f(k) (:306-310), which reads
// past the end of x1f whenever a block has more x3 faces than x1 faces; here
// x3f holds the x3 face coordinates.
#include "exo_netcdf.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace exo {

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int NG = 3;  // ghost cells on each side

enum NcType { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5 };

struct NcAtt {
  std::string name;
  int type;
  std::string text;
  std::vector<int32_t> ints;
};

struct NcVar {
  std::string name;
  std::vector<int> dims;        // dimension ids; dimension 0 is the record dimension
  std::vector<NcAtt> atts;
  std::vector<float> data;      // one record of a record variable, else all of it
};

struct NcFile {
  std::vector<std::pair<std::string, size_t>> dims;  // length 0: record dimension
  std::vector<NcAtt> gatts;
  std::vector<NcVar> vars;
};

NcAtt att_text(const char *name, const std::string &t) { return NcAtt{name, NC_CHAR, t, {}}; }
NcAtt att_ints(const char *name, std::vector<int32_t> v) {
  return NcAtt{name, NC_INT, "", std::move(v)};
}

// big-endian header serialisation
struct Bytes {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void atts(const std::vector<NcAtt> &a) {
    if (a.empty()) {
      u32(0);
      u32(0);
      return;
    }
    u32(0x0C);
    u32((uint32_t)a.size());
    for (auto &t : a) {
      name(t.name);
      u32((uint32_t)t.type);
      if (t.type == NC_CHAR) {
        u32((uint32_t)t.text.size());
        b.insert(b.end(), t.text.begin(), t.text.end());
      } else {
        u32((uint32_t)t.ints.size());
        for (int32_t v : t.ints) u32((uint32_t)v);
      }
      pad();
    }
  }
};

bool is_record(const NcFile &f, const NcVar &v) {
  return !v.dims.empty() && f.dims[v.dims[0]].second == 0;
}

size_t var_bytes(const NcFile &f, const NcVar &v) {
  size_t n = 4;
  for (size_t d = is_record(f, v) ? 1 : 0; d < v.dims.size(); ++d) n *= f.dims[v.dims[d]].second;
  return n;
}

std::vector<uint8_t> header(const NcFile &f, const std::vector<uint64_t> &begin) {
  Bytes h;
  h.b = {'C', 'D', 'F', 2};
  h.u32(1);  // numrecs
  h.u32(0x0A);
  h.u32((uint32_t)f.dims.size());
  for (auto &d : f.dims) {
    h.name(d.first);
    h.u32((uint32_t)d.second);
  }
  h.atts(f.gatts);
  h.u32(0x0B);
  h.u32((uint32_t)f.vars.size());
  for (size_t q = 0; q < f.vars.size(); ++q) {
    auto &v = f.vars[q];
    h.name(v.name);
    h.u32((uint32_t)v.dims.size());
    for (int d : v.dims) h.u32((uint32_t)d);
    h.atts(v.atts);
    h.u32(NC_FLOAT);
    size_t vs = var_bytes(f, v);
    h.u32(vs > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)vs);
    h.u64(begin[q]);
  }
  return h.b;
}

void put_floats(std::FILE *fp, const std::vector<float> &v, std::vector<uint8_t> &buf) {
  buf.resize(v.size() * 4);
  for (size_t e = 0; e < v.size(); ++e) {
    uint32_t x;
    std::memcpy(&x, &v[e], 4);
    buf[4 * e] = (uint8_t)(x >> 24);
    buf[4 * e + 1] = (uint8_t)(x >> 16);
    buf[4 * e + 2] = (uint8_t)(x >> 8);
    buf[4 * e + 3] = (uint8_t)x;
  }
  std::fwrite(buf.data(), 1, buf.size(), fp);
}

// Lays out and writes one file holding a single record.  Fixed-size variables
// follow the header in definition order, then the record of every record
// variable in definition order.
std::string write_nc3(const std::string &path, const NcFile &f) {
  std::vector<uint64_t> begin(f.vars.size(), 0);
  size_t hsize = header(f, begin).size();
  uint64_t off = hsize;
  for (size_t q = 0; q < f.vars.size(); ++q)
    if (!is_record(f, f.vars[q])) {
      begin[q] = off;
      off += var_bytes(f, f.vars[q]);
    }
  for (size_t q = 0; q < f.vars.size(); ++q)
    if (is_record(f, f.vars[q])) {
      begin[q] = off;
      off += var_bytes(f, f.vars[q]);
    }
  for (size_t q = 0; q < f.vars.size(); ++q)
    if (f.vars[q].data.size() * 4 != var_bytes(f, f.vars[q]))
      return "internal error: size of variable " + f.vars[q].name;
  std::FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) return "cannot open " + path + " for writing";
  std::vector<uint8_t> h = header(f, begin), buf;
  std::fwrite(h.data(), 1, h.size(), fp);
  for (int rec = 0; rec < 2; ++rec)
    for (auto &v : f.vars)
      if (is_record(f, v) == (rec == 1)) put_floats(fp, v.data, buf);
  bool bad = std::ferror(fp) != 0;
  if (std::fclose(fp) != 0) bad = true;
  if (bad) return "write error on " + path;
  return "";
}

// Athena's MeshBlockTree lists leaves depth first, children in (x3, x2, x1)
// bit order; on the cube the six panels are the level-2 nodes in order, so
// the block at (sub-row sk, sub-column sj) of panel p is number
// p * bps^2 + (Z-order of (sk, sj) with the x3 bit above the x2 bit).
int block_gid(int p, int sk, int sj, int bps) {
  int z = 0;
  for (int b = 0; (1 << b) < bps; ++b) {
    z |= ((sj >> b) & 1) << (2 * b);
    z |= ((sk >> b) & 1) << (2 * b + 1);
  }
  return p * bps * bps + z;
}

void units_of(const std::string &name, std::string *units, std::string *lname) {
  // output_utils.cpp:36-38 (the only entries that this path writes)
  if (name == "rho") {
    *units = "kg/m^3";
    *lname = "density";
  } else if (name == "press") {
    *units = "pa";
    *lname = "pressure";
  } else if (name.compare(0, 3, "vel") == 0) {
    *units = "m/s";
    *lname = "velocity";
  } else {
    units->clear();
    lname->clear();
  }
}

}  // namespace

int out_set_from_name(const char *v) {
  if (!v) return -1;
  if (!std::strcmp(v, "prim")) return OUT_PRIM;
  if (!std::strcmp(v, "cons")) return OUT_CONS;
  if (!std::strcmp(v, "uov")) return OUT_UOV;
  return -1;
}

std::vector<std::string> out_names(int set, int eos, int ntracer) {
  std::vector<std::string> n;
  if (set == OUT_UOV) return {"lat", "lon", "U", "V", "sqrtg"};
  bool ideal = eos == EXO_EOS_IDEAL;
  if (set == OUT_PRIM) {
    n.push_back("rho");
    if (ideal) n.push_back("press");
    for (int d = 1; d <= 3; ++d) n.push_back("vel" + std::to_string(d));
    for (int t = 0; t < ntracer; ++t) n.push_back("r" + std::to_string(t));
  } else {
    n.push_back("dens");
    if (ideal) n.push_back("Etot");
    for (int d = 1; d <= 3; ++d) n.push_back("mom" + std::to_string(d));
    for (int t = 0; t < ntracer; ++t) n.push_back("s" + std::to_string(t));
  }
  return n;
}

// MeshBlock::UserWorkBeforeOutput of examples/2023-Chen-exo3/W92.cpp:115-141:
// lat, lon (CubedSphere::GetLatLon, cubed_sphere.cpp:78-90) in degrees,
// sqrtg = delta / (C D), and the factors of CubedSphere::GetUV
// (cubed_sphere.cpp:132-145 over VecTransRLLFromABP,
// cubed_sphere_transform.cpp:36-61) grouped as the reference's expressions
// group them, so the device's products and sums round alike:
//   panels 2, 3, 4, 6: U = -(f1 V3),            V = f2 V3 - V2
//   panel 1:           U = -(f1 V2 + f2 V3),    V = f3 V2 - f4 V3
//   panel 5:           U = -(f1 V2 - f2 V3),    V = f3 V2 + f4 V3
std::vector<double> uov_table(const Setup &s) {
  size_t n2 = (size_t)6 * s.nc3 * s.nc2;
  std::vector<double> t(7 * n2, 0.);
  for (int p = 0; p < 6; ++p)
    for (int k = 0; k < s.nc3; ++k)
      for (int j = 0; j < s.nc2; ++j) {
        size_t c = ((size_t)p * s.nc3 + k) * s.nc2 + j;
        double lat, lon;
        s.latlon(p, k, j, &lat, &lon);
        t[0 * n2 + c] = lat / PI * 180.0;
        t[1 * n2 + c] = lon / PI * 180.0;
        double x = tan(s.x2v[p][j]), y = tan(s.x3v[p][k]);
        double C = sqrt(1.0 + x * x), D = sqrt(1.0 + y * y);
        double delta = 1.0 / (1.0 + x * x + y * y);
        t[2 * n2 + c] = delta / (C * D);
        double E = sqrt(x * x + y * y), dl = sqrt(1 + x * x + y * y);
        double f1, f2, f3 = 0., f4 = 0.;
        int b = p + 1;
        if (b == 1) {
          f1 = -dl * y / (D * E);
          f2 = dl * x / (C * E);
          f3 = -x / (D * E);
          f4 = y / (C * E);
        } else if (b == 5) {
          f1 = dl * y / (D * E);
          f2 = dl * x / (C * E);
          f3 = x / (D * E);
          f4 = y / (C * E);
        } else {
          f1 = dl / (C * D);
          f2 = x * y / (C * D);
        }
        t[3 * n2 + c] = f1;
        t[4 * n2 + c] = f2;
        t[5 * n2 + c] = f3;
        t[6 * n2 + c] = f4;
      }
  return t;
}

std::string write_output_files(const Setup &s, const std::string &basename, int set,
                               const std::string &file_id, int file_number, double time,
                               int bps, int ghost, const float *fields) {
  int N = s.N;
  if (bps < 1 || (bps & (bps - 1)) || N % bps)
    return "blocks_per_side must be a power of two dividing the panel size";
  if (file_number < 0 || file_number > 99999) return "file_number must be 0..99999";
  std::vector<std::string> names = out_names(set, s.cfg.eos, s.nt);
  int nvar = (int)names.size();
  int G = ghost ? NG : 0;
  int nb = N / bps;
  int n1 = s.nx1 + 2 * G, n2 = N + 2 * G, n3 = N + 2 * G;  // panel arrays
  int c1 = n1, c2 = nb + 2 * G, c3 = nb + 2 * G;            // one block
  int i0 = NG - G;                                          // first output x1 cell
  int nbtotal = 6 * bps * bps;
  char number[8];
  std::snprintf(number, sizeof(number), "%05d", file_number);
  for (int p = 0; p < 6; ++p)
    for (int sk = 0; sk < bps; ++sk)
      for (int sj = 0; sj < bps; ++sj) {
        int gid = block_gid(p, sk, sj, bps);
        int lx2 = (p % 2) * bps + sj, lx3 = (p / 2) * bps + sk;
        int j0 = NG - G + sj * nb, k0 = NG - G + sk * nb;  // first cell in Setup indices
        NcFile f;
        int dt = 0, d1, d1f = -1, d2, d2f = -1, d3, d3f = -1;
        f.dims.push_back({"time", 0});
        d1 = (int)f.dims.size();
        f.dims.push_back({"x1", (size_t)c1});
        if (c1 > 1) {
          d1f = (int)f.dims.size();
          f.dims.push_back({"x1f", (size_t)c1 + 1});
        }
        d2 = (int)f.dims.size();
        f.dims.push_back({"x2", (size_t)c2});
        if (c2 > 1) {
          d2f = (int)f.dims.size();
          f.dims.push_back({"x2f", (size_t)c2 + 1});
        }
        d3 = (int)f.dims.size();
        f.dims.push_back({"x3", (size_t)c3});
        if (c3 > 1) {
          d3f = (int)f.dims.size();
          f.dims.push_back({"x3f", (size_t)c3 + 1});
        }
        f.gatts.push_back(att_ints("NumFilesInSet", {nbtotal}));
        NcVar vt{"time", {dt}, {att_text("axis", "T"), att_text("units", "s"),
                                att_text("long_name", "time")}, {(float)time}};
        f.vars.push_back(vt);
        // coordinate axes: (dimension, face dimension, axis letter, mesh
        // cells, cells here, logical location, centres, faces)
        struct Ax {
          int d, df;
          const char *axis, *cname, *fname, *ctext, *ftext;
          int mesh_n, cells, loc;
          const std::vector<double> *v, *fc;
          int first;
        };
        Ax ax[3] = {
            {d1, d1f, "Z", "x1", "x1f", "Z-coordinate at cell center", "Z-coordinate at cell face",
             s.nx1, c1, 0, &s.x1v, &s.x1f, i0},
            {d2, d2f, "Y", "x2", "x2f", "Y-coordinate at cell center", "Y-coordinate at cell face",
             2 * N, c2, lx2, &s.x2v[p], &s.x2f[p], j0},
            {d3, d3f, "X", "x3", "x3f", "X-coordinate at cell center", "X-coordinate at cell face",
             3 * N, c3, lx3, &s.x3v[p], &s.x3f[p], k0}};
        for (auto &a : ax) {
          NcVar v{a.cname, {a.d}, {}, {}};
          v.atts = {att_text("axis", a.axis), att_text("units", "m"),
                    att_text("long_name", a.ctext),
                    att_ints("domain_decomposition",
                             {1, a.mesh_n, a.cells * a.loc + 1, a.cells * (a.loc + 1)})};
          for (int e = 0; e < a.cells; ++e) v.data.push_back((float)(*a.v)[a.first + e]);
          f.vars.push_back(v);
          if (a.df >= 0) {
            NcVar w{a.fname, {a.df}, {}, {}};
            w.atts = {att_text("units", "m"), att_text("long_name", a.ftext),
                      att_ints("domain_decomposition",
                               {0, a.mesh_n, a.cells * a.loc, a.cells * (a.loc + 1)})};
            for (int e = 0; e <= a.cells; ++e) w.data.push_back((float)(*a.fc)[a.first + e]);
            f.vars.push_back(w);
          }
        }
        for (int q = 0; q < nvar; ++q) {
          NcVar v{names[q], {dt, d1, d2, d3}, {}, {}};
          std::string units, lname;
          units_of(names[q], &units, &lname);
          if (!units.empty()) v.atts.push_back(att_text("units", units));
          if (!lname.empty()) v.atts.push_back(att_text("long_name", lname));
          v.data.resize((size_t)c1 * c2 * c3);
          const float *src = fields + ((size_t)p * nvar + q) * n1 * n2 * n3;
          size_t e = 0;
          for (int i = 0; i < c1; ++i)
            for (int j = 0; j < c2; ++j) {
              const float *row = src + ((size_t)i * n2 + (sj * nb + j)) * n3 + sk * nb;
              for (int k = 0; k < c3; ++k) v.data[e++] = row[k];
            }
          f.vars.push_back(std::move(v));
        }
        std::string path = basename + ".block" + std::to_string(gid) + "." + file_id + "." +
                           number + ".nc";
        std::string m = write_nc3(path, f);
        if (!m.empty()) return m;
      }
  return "";
}

}  // namespace exo
