// test_host.cpp -- the C++ host interface (include/exocubed.hpp) exercised the
// way the reference's own tests exercise its classes (tests/test_weno.cpp).
// "host" runs the checks that need no GPU; "all" adds the device checks.
// Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "exocubed.hpp"

namespace {

struct Case {
  const char *suite, *name;
  bool device;
  std::function<void()> fn;
};
std::vector<Case> &registry() {
  static std::vector<Case> r;
  return r;
}
int g_failures = 0;

struct Register {
  Register(const char *s, const char *n, bool dev, std::function<void()> f) {
    registry().push_back({s, n, dev, std::move(f)});
  }
};

#define TEST_CASE(suite, name, dev)                                              \
  void suite##_##name();                                                         \
  Register reg_##suite##_##name(#suite, #name, dev, suite##_##name);             \
  void suite##_##name()
#define HOST_TEST(suite, name) TEST_CASE(suite, name, false)
#define DEVICE_TEST(suite, name) TEST_CASE(suite, name, true)

#define EXPECT_TRUE(c)                                                            \
  do {                                                                            \
    if (!(c)) {                                                                   \
      std::printf("  %s:%d: expected %s\n", __FILE__, __LINE__, #c);              \
      ++g_failures;                                                               \
    }                                                                             \
  } while (0)
#define EXPECT_NEAR(a, b, tol)                                                    \
  do {                                                                            \
    double a_ = (a), b_ = (b);                                                    \
    if (!(std::fabs(a_ - b_) <= (tol))) {                                         \
      std::printf("  %s:%d: %s = %.17g, expected %.17g\n", __FILE__, __LINE__, #a, a_, b_); \
      ++g_failures;                                                               \
    }                                                                             \
  } while (0)

const double R_EARTH = 6371000.0;

exo_config sw_config(int nx, int nx1, int rsolver) {
  exo_config c{};
  c.nx = nx;
  c.nx1 = nx1;
  c.eos = EXO_EOS_SHALLOW_YZ;
  c.rsolver = rsolver;
  c.coriolis = 1;
  c.x1min = R_EARTH;
  c.x1max = R_EARTH + 1.0;
  c.gamma = 1.4;
  c.omega = 7.292e-5;
  c.cfl = 0.4;
  c.precision = EXO_PREC_FP64;
  return c;
}

exo_config atm_config(int nx, int nx1, double cfl) {
  exo_config c{};
  c.nx = nx;
  c.nx1 = nx1;
  c.eos = EXO_EOS_IDEAL;
  c.rsolver = 3;  // LMARS
  c.implicit_flag = 9;
  c.cfl_x1 = 1;
  c.x1min = R_EARTH;
  c.x1max = R_EARTH + 30000.0;
  c.gamma = 1.4;
  c.grav = -9.81;
  c.omega = 7.292e-5;
  c.cfl = cfl;
  c.precision = EXO_PREC_FP64;
  return c;
}

bool file_starts_with(const std::string &path, const char *magic, size_t n) {
  std::FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  char b[8] = {0};
  size_t got = std::fread(b, 1, n, f);
  std::fclose(f);
  return got == n && std::memcmp(b, magic, n) == 0;
}

std::string tmpdir() {
  char tmpl[] = "/tmp/exo_host_XXXXXX";
  const char *d = mkdtemp(tmpl);
  return d ? d : "/tmp";
}

// ---------------------------------------------------------------------------
HOST_TEST(output, variable_names) {
  exo_config c = sw_config(8, 3, 0);
  char names[256];
  int n = 0;
  EXPECT_TRUE(exo_output_names(&c, "prim", names, sizeof(names), &n) == EXO_OK);
  EXPECT_TRUE(n == 4 && std::string(names) == "rho,vel1,vel2,vel3");
  EXPECT_TRUE(exo_output_names(&c, "uov", names, sizeof(names), &n) == EXO_OK);
  EXPECT_TRUE(std::string(names) == "lat,lon,U,V,sqrtg");
  EXPECT_TRUE(exo_output_names(&c, "radiance", names, sizeof(names), &n) == EXO_ERR_ARG);
}

HOST_TEST(gnomonic_equiangle, metric_identity) {
  // cos^2 + sin^2 = 1 of the cell metric on every panel
  exo_config c = sw_config(8, 3, 0);
  size_t n = 0;
  EXPECT_TRUE(exo_describe(&c, 2, nullptr, 0, &n) == EXO_OK);
  std::vector<double> tab(n);
  EXPECT_TRUE(exo_describe(&c, 2, tab.data(), n * sizeof(double), &n) == EXO_OK);
  size_t plane = n / 20;  // 20 tables of [6][nc3][nc2]
  double worst = 0;
  for (size_t q = 0; q < plane; ++q)
    worst = std::fmax(worst, std::fabs(tab[q] * tab[q] + tab[plane + q] * tab[plane + q] - 1.0));
  EXPECT_NEAR(worst, 0.0, 1e-15);
  EXPECT_TRUE(exo_footprint(&c) > 0);
  c.nx = 2;  // below the smallest panel
  EXPECT_TRUE(exo_footprint(&c) == 0);
}

HOST_TEST(netcdf, files_from_host_fields) {
  exo_config c = sw_config(8, 3, 0);
  std::vector<float> f((size_t)6 * 4 * 3 * 8 * 8, 1.f);
  std::string base = tmpdir() + "/W92";
  EXPECT_TRUE(exo_write_netcdf_fields(&c, base.c_str(), "prim", nullptr, 2, 60.0, 2, 0, f.data()) ==
              EXO_OK);
  exo::NetcdfOutput out(base, "prim", 60.0, 2);
  for (int gid = 0; gid < 24; ++gid)
    EXPECT_TRUE(file_starts_with(out.FileName(gid, 2), "CDF\x02", 4));
  EXPECT_TRUE(exo_write_netcdf_fields(&c, "/nonexistent/dir/x", "prim", nullptr, 0, 0.0, 1, 0,
                                      f.data()) == EXO_ERR_IO);
}

// ---------------------------------------------------------------------------
DEVICE_TEST(interp_weno5, test_case1) {
  // tests/test_weno.cpp:20-29 of the reference, on the device
  exo::Instance in(sw_config(8, 3, 0));
  double in5[5] = {1.0, 2.0, 3.0, 4.0, 5.0}, out = 0;
  EXPECT_TRUE(exo_weno5_batch(in.handle(), in5, &out, 1) == EXO_OK);
  EXPECT_NEAR(out, 2.5000000000000004, 1.E-10);
  EXPECT_TRUE(out == 2.5000000000000004);
}

DEVICE_TEST(hydro, stage_by_parts_equals_stage) {
  exo_config c = sw_config(12, 3, 1);
  exo::Instance a(c), b(c);
  std::vector<double> u0 = a.InitialState(0, 20261202, {});
  a.Upload(u0);
  b.Upload(u0);
  exo::Hydro hydro(a);
  exo::BoundaryValues bv(a);
  double dt = hydro.NewBlockTimeStep();
  EXPECT_TRUE(dt > 0);
  hydro.CalculateFluxes();
  hydro.AddFluxDivergence(dt);
  exo::ImplicitSolver(a).SolveImplicit3D(dt);  // shallow water: nothing to solve
  bv.ApplyPhysicalBoundaries();
  exo::Mesh(b).DoTaskListOneStage(1, dt);
  std::vector<double> ua = a.Download(), ub = b.Download();
  int nc1 = c.nx1 + 6, nc = c.nx + 6, nv = a.nvar();
  size_t diff = 0;
  for (int p = 0; p < 6; ++p)
    for (int n = 0; n < nv; ++n)
      for (int k = 3; k < nc - 3; ++k)
        for (int j = 3; j < nc - 3; ++j)
          for (int i = 3; i < 3 + c.nx1; ++i) {
            size_t q = ((((size_t)p * nv + n) * nc + k) * nc + j) * nc1 + i;
            diff += ua[q] != ub[q];
          }
  EXPECT_TRUE(diff == 0);
}

DEVICE_TEST(implicit_solver, partial_correction_by_parts) {
  // implicit_flag 1 (PartialCorrection): the stage through the classes equals
  // the whole stage, and the 3x3 solve changes the state
  exo_config c = atm_config(6, 8, 0.4);
  c.implicit_flag = 1;
  exo::Instance a(c), b(c);
  std::vector<double> u0 = a.InitialState(1, 20261202, {});
  a.Upload(u0);
  b.Upload(u0);
  exo::Hydro hydro(a);
  double dt = hydro.NewBlockTimeStep();
  hydro.CalculateFluxes();
  hydro.AddFluxDivergence(dt);
  exo::ImplicitSolver(a).SolveImplicit3D(dt);
  exo::BoundaryValues(a).ApplyPhysicalBoundaries();
  exo::Mesh(b).DoTaskListOneStage(1, dt);
  std::vector<double> ua = a.Download(), ub = b.Download();
  EXPECT_TRUE(ua == ub);
  EXPECT_TRUE(ua != u0);
  // stages 2 and 3 by parts: IntegrateHydro forms the stage's weighted average
  const double beta[3] = {1.0, 0.25, 2.0 / 3.0};
  for (int stage = 2; stage <= 3; ++stage) {
    hydro.CalculateFluxes();
    hydro.IntegrateHydro(stage, dt);
    exo::ImplicitSolver(a).SolveImplicit3D(beta[stage - 1] * dt);
    exo::BoundaryValues(a).ApplyPhysicalBoundaries();
    exo::Mesh(b).DoTaskListOneStage(stage, dt);
    EXPECT_TRUE(a.Download() == b.Download());
  }
}

DEVICE_TEST(mesh, redo_then_checkpoint_restart) {
  exo_config c = atm_config(6, 6, 1.5);  // a CFL the first steps do not survive
  exo::Instance a(c);
  a.Upload(a.InitialState(1, 20261202, {}));
  exo::Mesh mesh(a);
  double dt = 0;
  int redos = mesh.StepWithRedo(3, 5, &dt);
  EXPECT_TRUE(redos > 0 && dt > 0);
  EXPECT_TRUE(mesh.CheckAllValid());
  std::string path = tmpdir() + "/run.ckpt";
  mesh.SaveCheckpoint(path, 123.0, 3);
  mesh.Step(2);
  exo::Instance b(c);
  exo::Mesh mb(b);
  std::pair<double, int64_t> tn = mb.LoadCheckpoint(path);
  EXPECT_TRUE(tn.first == 123.0 && tn.second == 3);
  mb.Step(2);
  EXPECT_TRUE(a.Download() == b.Download());
  bool threw = false;
  try {
    mb.LoadCheckpoint(path + ".missing");
  } catch (const exo::RuntimeError &e) {
    threw = std::strstr(e.what(), "cannot open") != nullptr;
  }
  EXPECT_TRUE(threw);
}

DEVICE_TEST(netcdf_output, cadence) {
  exo::Instance in(sw_config(8, 3, 0));
  in.Upload(in.InitialState(0, 7, {}));
  std::string base = tmpdir() + "/W92";
  exo::NetcdfOutput out(base, "uov", 100.0);
  int written = 0;
  for (double t : {0.0, 50.0, 100.0, 150.0}) written += out.MakeOutputs(in, t);
  EXPECT_TRUE(written == 2 && out.file_number() == 2);
  for (int gid = 0; gid < 6; ++gid) {
    EXPECT_TRUE(file_starts_with(out.FileName(gid, 0), "CDF\x02", 4));
    EXPECT_TRUE(file_starts_with(out.FileName(gid, 1), "CDF\x02", 4));
  }
}

}  // namespace

int main(int argc, char **argv) {
  bool device = argc > 1 && std::string(argv[1]) == "all";
  int run = 0;
  for (auto &c : registry()) {
    if (c.device && !device) continue;
    std::printf("[ RUN      ] %s.%s\n", c.suite, c.name);
    int before = g_failures;
    try {
      c.fn();
    } catch (const std::exception &e) {
      std::printf("  exception: %s\n", e.what());
      ++g_failures;
    }
    std::printf("[ %s ] %s.%s\n", g_failures == before ? "      OK" : " FAILED ", c.suite,
                c.name);
    ++run;
  }
  std::printf("%d tests, %d failed checks\n", run, g_failures);
  return g_failures == 0 ? 0 : 1;
}
