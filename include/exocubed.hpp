// exocubed.hpp -- C++ host interface over the C ABI of exocubed.h.
//
// The classes are named after the reference's (Athena++ / canoe) classes
// whose operations they replace, so a host written against the reference's
// call surface calls the same operations on the MI355X instance:
//   Hydro           Hydro::CalculateFluxes (patches/19.calculate_fluxes.patch:
//                   17-163), Hydro::AddFluxDivergence with
//                   GnomonicEquiangle::AddCoordTermsDivergence
//                   (gnomonic_equiangle.cpp:834-879) and the sources,
//                   Hydro::NewBlockTimeStep (patches/21.new_blockdt.patch)
//   ImplicitSolver  ImplicitSolver::SolveImplicit3D (solve_implicit_3d.cpp:
//                   27-180) with ImplicitHydroTasks::ImplicitCorrection
//   BoundaryValues  the cubed-sphere ghost exchange (cubed_sphere_utility.cpp:
//                   371-617), then cons->prim and the reflecting x1 boundary
//   Mesh            the main loop's step, time step and redo path
//                   (main.cpp.in:126-186, src/mesh_redo.cpp:12-45), restart
//   NetcdfOutput    NetcdfOutput::WriteOutputFile (src/outputs/netcdf.cpp)
// One Instance holds the six panels of one model on the device (an
// exo_handle).  Every failure throws exo::RuntimeError with the library's
// message, as the reference throws RuntimeError / ATHENA_ERROR.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "exocubed.h"

namespace exo {

using Real = double;

class RuntimeError : public std::runtime_error {
 public:
  RuntimeError(const std::string &where, const std::string &msg)
      : std::runtime_error(where + ": " + msg) {}
};

inline void check(int rc, const exo_handle *h, const char *where) {
  if (rc == EXO_OK) return;
  std::string msg = h ? exo_last_error(h) : "";
  if (msg.empty()) msg = "error code " + std::to_string(rc);
  throw RuntimeError(where, msg);
}

// The state of one instance on the device.  Host arrays use Athena's
// per-MeshBlock layout, one block per panel: (6, NVAR, nc3, nc2, nc1).
class Instance {
 public:
  explicit Instance(const exo_config &cfg) : cfg_(cfg) {
    check(exo_create(&cfg_, &h_), nullptr, "exo_create");
  }
  ~Instance() {
    if (h_) exo_destroy(h_);
  }
  Instance(const Instance &) = delete;
  Instance &operator=(const Instance &) = delete;

  exo_handle *handle() const { return h_; }
  const exo_config &config() const { return cfg_; }
  int nvar() const { return exo_nvar(h_); }
  size_t state_size() const { return exo_state_doubles(h_); }
  size_t scalar_size() const {
    return (size_t)6 * cfg_.ntracer * (cfg_.nx + 6) * (cfg_.nx + 6) * (cfg_.nx1 + 6);
  }

  // the seeded synthetic initial states of the benchmark configurations
  std::vector<Real> InitialState(int kind, uint64_t seed, const std::vector<Real> &params,
                                 std::vector<Real> *scalars = nullptr) const {
    std::vector<Real> u(state_size()), s(scalar_size());
    check(exo_init_state(h_, kind, seed, params.empty() ? nullptr : params.data(),
                         (int)params.size(), u.data(), s.empty() ? nullptr : s.data()),
          h_, "exo_init_state");
    if (scalars) *scalars = std::move(s);
    return u;
  }
  void Upload(const std::vector<Real> &u, const std::vector<Real> *s = nullptr) {
    check(exo_upload_state(h_, u.data(), s ? s->data() : nullptr), h_, "exo_upload_state");
  }
  std::vector<Real> Download(std::vector<Real> *s = nullptr) const {
    std::vector<Real> u(state_size());
    if (s) s->resize(scalar_size());
    check(exo_download_state(h_, u.data(), s ? s->data() : nullptr), h_, "exo_download_state");
    return u;
  }
  std::vector<Real> Primitives() const {
    std::vector<Real> w(state_size());
    check(exo_download_prim(h_, w.data()), h_, "exo_download_prim");
    return w;
  }

 private:
  exo_config cfg_;
  exo_handle *h_ = nullptr;
};

class Hydro {
 public:
  explicit Hydro(Instance &in) : h_(in.handle()), in_(in) {}
  // interface fluxes of every panel: reconstruction, decomposition, edge
  // exchange and Riemann solver
  void CalculateFluxes() { check(exo_calc_fluxes(h_), h_, "Hydro::CalculateFluxes"); }
  // flux[dir] in Athena's flux layout (interior faces defined)
  std::vector<Real> Flux(int dir) const {
    const exo_config &c = in_.config();
    size_t n1 = c.nx1 + 6 + (dir == 0), n2 = c.nx + 6 + (dir == 1), n3 = c.nx + 6 + (dir == 2);
    std::vector<Real> f((size_t)6 * in_.nvar() * n1 * n2 * n3);
    check(exo_download_flux(h_, dir, f.data()), h_, "Hydro::Flux");
    return f;
  }
  // -div F + coordinate and physical sources times wght (stage-1 weights):
  // into u for shallow water, into the tendency du for the ideal gas
  void AddFluxDivergence(Real wght) {
    check(exo_add_flux_divergence(h_, wght), h_, "Hydro::AddFluxDivergence");
  }
  // ImplicitHydroTasks::IntegrateHydro + AddSourceTerms of SSPRK3 stage 1..3:
  // the stage's weighted average of u, then the update with beta(stage) dt
  void IntegrateHydro(int stage, Real dt) {
    check(exo_integrate_hydro(h_, stage, dt), h_, "Hydro::IntegrateHydro");
  }
  // the CFL time step of the current state
  Real NewBlockTimeStep() {
    Real dt = 0;
    check(exo_new_dt(h_, &dt), h_, "Hydro::NewBlockTimeStep");
    return dt;
  }

 private:
  exo_handle *h_;
  Instance &in_;
};

class ImplicitSolver {
 public:
  explicit ImplicitSolver(Instance &in) : h_(in.handle()) {}
  // the vertical block-tridiagonal correction of du (5x5 with implicit_flag 9,
  // 3x3 with 1), then u += du
  void SolveImplicit3D(Real dt_stage) {
    check(exo_implicit_correct(h_, dt_stage), h_, "ImplicitSolver::SolveImplicit3D");
  }

 private:
  exo_handle *h_;
};

class BoundaryValues {
 public:
  explicit BoundaryValues(Instance &in) : h_(in.handle()) {}
  void ApplyPhysicalBoundaries() {
    check(exo_fill_ghosts(h_), h_, "BoundaryValues::ApplyPhysicalBoundaries");
  }

 private:
  exo_handle *h_;
};

class Mesh {
 public:
  explicit Mesh(Instance &in) : h_(in.handle()) {}
  // one SSPRK3 stage with the step's dt (the task list's stage)
  void DoTaskListOneStage(int stage, Real dt) {
    check(exo_stage(h_, stage, dt), h_, "Mesh::DoTaskListOneStage");
  }
  // nsteps SSPRK3 steps, each at the CFL dt (dt_fixed <= 0) or dt_fixed;
  // returns the dt of the last step
  Real Step(int nsteps, Real dt_fixed = 0) {
    Real dt = 0;
    check(exo_step(h_, nsteps, dt_fixed, &dt), h_, "Mesh::Step");
    return dt;
  }
  void SaveAllStates() { check(exo_save_states(h_), h_, "Mesh::SaveAllStates"); }
  void LoadAllStates() { check(exo_load_states(h_), h_, "Mesh::LoadAllStates"); }
  bool CheckAllValid() {
    int v = 0;
    check(exo_state_valid(h_, &v), h_, "Mesh::CheckAllValid");
    return v != 0;
  }
  // the main loop with its redo path; returns the steps redone
  int StepWithRedo(int nsteps, int max_redo, Real *dt_last = nullptr) {
    int redos = 0;
    check(exo_step_redo(h_, nsteps, max_redo, dt_last, &redos), h_, "Mesh::StepWithRedo");
    return redos;
  }
  std::pair<Real, Real> TimeStepState() {
    Real dt = 0, prev = 0;
    check(exo_get_dt_state(h_, &dt, &prev), h_, "Mesh::TimeStepState");
    return {dt, prev};
  }
  void SetTimeStepState(Real dt, Real dt_prev) {
    check(exo_set_dt_state(h_, dt, dt_prev), h_, "Mesh::SetTimeStepState");
  }
  void SaveCheckpoint(const std::string &path, Real time, int64_t ncycle) {
    check(exo_save_checkpoint(h_, path.c_str(), time, ncycle), h_, "Mesh::SaveCheckpoint");
  }
  // restores a checkpoint; returns (time, ncycle)
  std::pair<Real, int64_t> LoadCheckpoint(const std::string &path) {
    Real t = 0;
    int64_t n = 0;
    check(exo_load_checkpoint(h_, path.c_str(), &t, &n), h_, "Mesh::LoadCheckpoint");
    return {t, n};
  }

 private:
  exo_handle *h_;
};

// One <outputN> block of file_type netcdf: writes every dt of simulated time
// (file_number and next_time advance as netcdf.cpp:459-466 advances them).
class NetcdfOutput {
 public:
  NetcdfOutput(std::string basename, std::string variable, Real dt, int blocks_per_side = 1,
               bool ghost_zones = false, std::string file_id = "")
      : basename_(std::move(basename)), variable_(std::move(variable)),
        file_id_(file_id.empty() ? variable_ : std::move(file_id)), dt_(dt),
        bps_(blocks_per_side), ghost_(ghost_zones) {}
  void WriteOutputFile(Instance &in, Real time) {
    check(exo_write_netcdf(in.handle(), basename_.c_str(), variable_.c_str(), file_id_.c_str(),
                           file_number_, time, bps_, ghost_ ? 1 : 0),
          in.handle(), "NetcdfOutput::WriteOutputFile");
    ++file_number_;
    next_time_ += dt_;
  }
  bool MakeOutputs(Instance &in, Real time) {
    if (time < next_time_) return false;
    WriteOutputFile(in, time);
    return true;
  }
  int file_number() const { return file_number_; }
  // the name of block gid's file of output number n
  std::string FileName(int gid, int n) const {
    char num[8];
    std::snprintf(num, sizeof(num), "%05d", n);
    return basename_ + ".block" + std::to_string(gid) + "." + file_id_ + "." + num + ".nc";
  }

 private:
  std::string basename_, variable_, file_id_;
  Real dt_;
  int bps_;
  bool ghost_;
  int file_number_ = 0;
  Real next_time_ = 0;
};

}  // namespace exo
