// Reference-shaped C++ operator API of the MI355X block-LB solver (header-only, over lbm_capi.h).
//
// Mirrors the call surface a user of the Peano block-LB drives (paths relative to
// madpac/peano/applications/latticeboltzmann/, BLB/ = blocklatticeboltzmann/):
//   BLB/configurations/BlockLatticeBoltzmannBatchJobConfigurationForRegularGrid  -> Configuration
//   BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid::run/runAsMaster -> Runner
//   BLB/repositories/...RepositoryForRegularGrid{,StandardImplementation}
//        switchToRegularBlockSolverAdapter / switchToRegularBlockSolverAndVTKOutputAdapter /
//        iterate / getRegularGridState / terminate / logIterationStatistics       -> Repository
//   BLB/RegularGridBlockState (tau, viscosity_L, dx, dt, timesteps, cell counter)     -> RegularGridBlockState
//   BLB/mappings/RegularGrid2RegularBlockSolver begin/endIteration                   -> RegularGrid2RegularBlockSolver
//   BLB/services/GridManagementService loadVertex/getPdf/getInner/getBoundary/
//        getDensity/getVelocity (6^3 blocks, vertex numbers from 8)                  -> BlockView via Repository::loadVertex
//   LB/PerformanceMeasurement (MLUPS per iteration; wall clock here, not CPU time)
// The three sweeps of one reference iteration (per-block collide + local streaming, the
// cross-block streaming of handleCell, the in-cell swap) are one fused device step here, issued at
// endIteration.  Errors that the reference reports through assertions throw std::runtime_error.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_capi.h"

namespace madpac_amd {
namespace blocklatticeboltzmann {

constexpr int LB_BLOCKSIZE = 6;                                    // LB/UserInput.h:24
constexpr int LB_BLOCK_NUMBER_OF_CELLS = 216;                      // BLB/BlockDefinitions.h:37
constexpr int LB_CURRENT_DIR = 19;

// lid-driven-cavity and default-channel of SLB/configurations/BlockLatticeBoltzmannScenarioConfiguration.cpp:52-56;
// PeriodicBox (all axes periodic, body force) has no reference counterpart
enum class Scenario { LidDrivenCavity, DefaultChannel, PeriodicBox };
enum class Precision { FP64 = LBM_FP64, FP32Shifted = LBM_FP32_SHIFTED, FP32ShiftedF32Math = LBM_FP32_SHIFTED_F32 };

// Run configuration: the XML tags regular-grid, lattice-boltzmann-configuration and
// lattice-boltzmann-scenario of the reference (LB/configurations/LatticeBoltzmannConfiguration.cpp:3-14,
// SLB configurations :3-11) as plain fields.
struct Configuration {
  int numberOfGridPoints[3] = {64, 64, 64};  // LB cells per dimension ("number-of-vertices")
  double domainSize[3] = {1.0, 1.0, 1.0};
  double tau = -1.0;                         // give tau or viscosity_L (tau wins when both are set)
  double viscosityL = -1.0;
  double dt = 1.0;
  double density = 1.0;
  int numberOfTimesteps = 1000;
  bool measurePerformance = false;
  Scenario scenario = Scenario::LidDrivenCavity;
  double characteristicVelocityL = 0.05;     // lid velocity (cavity) or mean inlet velocity (channel)
  bool useParabolicProfile = false;          // channel inlet profile ("use-parabolic-profile")
  double bodyForce[3] = {0.0, 0.0, 0.0};     // Guo forcing (periodic box)
  Precision precision = Precision::FP64;
  int device = 0;

  // LB/configurations/LatticeBoltzmannConfiguration.cpp:247-269: tau <-> viscosity_L
  void computeUndefinedValues() {
    if (tau > 0.0) {
      viscosityL = 1.0 / 3.0 * (tau - 0.5);
    } else if (viscosityL > 0.0) {
      tau = 3.0 * viscosityL + 0.5;
      if (tau < 0.5001 || tau > 1.9999) throw std::runtime_error("'tau' is out of range (0.5001, 1.9999)");
    } else {
      throw std::runtime_error("neither 'tau' nor 'viscosity_L' is given");
    }
  }
  double getDx() const { return domainSize[0] / numberOfGridPoints[0]; }
};

// BLB/RegularGridBlockState: the per-run state record
class RegularGridBlockState {
 public:
  double getTau() const { return tau_; }
  double getViscosityL() const { return viscosityL_; }
  double getDx() const { return dx_; }
  double getDt() const { return dt_; }
  double getDensity() const { return density_; }
  int getNumberTimesteps() const { return numberTimesteps_; }
  int getTimestep() const { return timestep_; }
  int64_t getLocalNumberOfCells() const { return cells_; }   // inner cells updated last iteration
  bool measurePerformance() const { return measure_; }
  double getLastMLUPS() const { return mlups_; }

 private:
  friend class Repository;
  double tau_ = 0, viscosityL_ = 0, dx_ = 0, dt_ = 0, density_ = 1, mlups_ = 0;
  int numberTimesteps_ = 0, timestep_ = 0;
  int64_t cells_ = 0;
  bool measure_ = false;
};

// One 6^3 block as GridManagementService holds it (BLB/services/GridManagementService.h:74-86):
// AoS pdf[cell*19 + q], cell = x + 6y + 36z.
struct BlockView {
  unsigned vertexNumber = 0;
  std::vector<uint8_t> inner, boundary;   // 216 flags
  std::vector<double> pdf;                // 216 * 19
  std::vector<double> density;            // 216
  std::vector<double> velocity;           // 216 * 3
};

class Repository {
 public:
  explicit Repository(const Configuration& cfg) : cfg_(cfg) {
    cfg_.computeUndefinedValues();
    lbm_config c{};
    c.nx = cfg_.numberOfGridPoints[0];
    c.ny = cfg_.numberOfGridPoints[1];
    c.nz = cfg_.numberOfGridPoints[2];
    const bool box = cfg_.scenario == Scenario::PeriodicBox;
    for (int d = 0; d < 3; d++) {
      c.periodic[d] = box ? 1 : 0;
      c.force[d] = box ? cfg_.bodyForce[d] : 0.0;
    }
    c.precision = static_cast<int>(cfg_.precision);
    c.tau = cfg_.tau;
    c.device = cfg_.device;
    check(lbm_create(&ctx_, &c), nullptr, "lbm_create");
    n_ = (int64_t)c.nx * c.ny * c.nz;
    // every cell of the box is fluid: the walls, lid, inlet and outlet lie outside the box
    // (GridManagementService::createVertex flags, BLB/services/GridManagementService.cpp:91-111)
    std::vector<uint8_t> inner((size_t)n_, 1);
    check(lbm_set_flags(ctx_, inner.data()), ctx_, "lbm_set_flags");
    if (cfg_.scenario == Scenario::LidDrivenCavity)
      check(lbm_set_cavity_lid(ctx_, cfg_.characteristicVelocityL), ctx_, "lbm_set_cavity_lid");
    else if (cfg_.scenario == Scenario::DefaultChannel)
      check(lbm_set_channel(ctx_, cfg_.characteristicVelocityL, cfg_.useParabolicProfile ? 1 : 0), ctx_,
            "lbm_set_channel");
    // f = w, rho = 1, u = 0 (GridManagementService.cpp:121-148)
    check(lbm_init_equilibrium(ctx_, nullptr, nullptr), ctx_, "lbm_init_equilibrium");
    state_.tau_ = lbm_tau_eff(cfg_.tau);
    state_.viscosityL_ = cfg_.viscosityL;
    state_.dx_ = cfg_.getDx();
    state_.dt_ = cfg_.dt;
    state_.density_ = cfg_.density;
    state_.numberTimesteps_ = cfg_.numberOfTimesteps;
    state_.measure_ = cfg_.measurePerformance;
  }
  Repository(const Repository&) = delete;
  Repository& operator=(const Repository&) = delete;
  ~Repository() { terminate(); }

  void switchToRegularBlockSolverAdapter() { vtk_ = false; }
  void switchToRegularBlockSolverAndVTKOutputAdapter(const std::string& basename) {
    vtk_ = true;
    vtkBase_ = basename;
  }

  // one time step: beginIteration, the fused A/B/C sweeps, endIteration
  void iterate() {
    const auto t0 = std::chrono::steady_clock::now();
    check(lbm_step(ctx_, 1), ctx_, "lbm_step");
    check(lbm_synchronize(ctx_), ctx_, "lbm_synchronize");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    state_.cells_ = lbm_fluid_cells(ctx_);
    state_.mlups_ = s > 0 ? (double)state_.cells_ / (1e6 * s) : 0.0;
    totalSeconds_ += s;
    iterations_++;
    state_.timestep_++;
    if (vtk_) writeVTK(vtkBase_ + "_" + std::to_string(state_.timestep_) + ".vtk");
  }

  RegularGridBlockState& getRegularGridState() { return state_; }

  void terminate() {
    if (ctx_) lbm_destroy(ctx_);
    ctx_ = nullptr;
  }

  void logIterationStatistics() const {
    std::printf("|| adapter name || iterations || total wall time [s] || average wall time [s]\n");
    std::printf("| RegularBlockSolverAdapter | %ld | %g | %g\n", (long)iterations_, totalSeconds_,
                iterations_ ? totalSeconds_ / iterations_ : 0.0);
  }

  // ---- block store view (GridManagementService) ----
  // The reference places vertices dx*6 apart starting at the domain corner and gives each a 6^3
  // block centred on it, blocks(d) = N/6 + 1 (+1 if N%6) per dimension
  // (BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:78-92); block slot s holds
  // the cell centred at (s - 2.5) dx, i.e. lattice cell s - 3.  Vertex numbers start at 8 (0..7 are
  // the temporary vertices, GridManagementService.cpp:26-54) in lexicographic block order.
  int blocksPerDimension(int d) const {
    const int n = cfg_.numberOfGridPoints[d];
    return n / LB_BLOCKSIZE + 1 + (n % LB_BLOCKSIZE ? 1 : 0);
  }
  unsigned numberOfVertices() const {
    return (unsigned)(blocksPerDimension(0) * blocksPerDimension(1) * blocksPerDimension(2));
  }
  BlockView loadVertex(unsigned vertexNumber) {
    if (vertexNumber < 8 || vertexNumber >= 8 + numberOfVertices())
      throw std::runtime_error("Could not find vertex " + std::to_string(vertexNumber));
    refreshHostCopy();
    const int b0 = blocksPerDimension(0), b1 = blocksPerDimension(1);
    const unsigned k = vertexNumber - 8;
    const int bx = (int)(k % b0), by = (int)((k / b0) % b1), bz = (int)(k / (b0 * b1));
    const int nx = cfg_.numberOfGridPoints[0], ny = cfg_.numberOfGridPoints[1], nz = cfg_.numberOfGridPoints[2];
    BlockView v;
    v.vertexNumber = vertexNumber;
    v.inner.assign(LB_BLOCK_NUMBER_OF_CELLS, 0);
    v.boundary.assign(LB_BLOCK_NUMBER_OF_CELLS, 0);
    v.pdf.assign(LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, 0.0);
    v.density.assign(LB_BLOCK_NUMBER_OF_CELLS, 0.0);
    v.velocity.assign(LB_BLOCK_NUMBER_OF_CELLS * 3, 0.0);
    for (int z = 0; z < LB_BLOCKSIZE; z++)
      for (int y = 0; y < LB_BLOCKSIZE; y++)
        for (int x = 0; x < LB_BLOCKSIZE; x++) {
          const int cell = x + LB_BLOCKSIZE * (y + LB_BLOCKSIZE * z);
          const int gx = 6 * bx + x - 3, gy = 6 * by + y - 3, gz = 6 * bz + z - 3;
          if (gx < 0 || gy < 0 || gz < 0 || gx >= nx || gy >= ny || gz >= nz) {
            // outside the closed domain: not inner; "boundary" when it touches the box
            const bool near = gx >= -1 && gy >= -1 && gz >= -1 && gx <= nx && gy <= ny && gz <= nz;
            v.boundary[cell] = near ? 1 : 0;
            continue;
          }
          const int64_t i = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
          v.inner[cell] = 1;
          v.boundary[cell] = (gx == 0 || gy == 0 || gz == 0 || gx == nx - 1 || gy == ny - 1 || gz == nz - 1) ? 1 : 0;
          for (int q = 0; q < LB_CURRENT_DIR; q++) v.pdf[cell * LB_CURRENT_DIR + q] = f_[(size_t)q * n_ + i];
          v.density[cell] = rho_[i];
          for (int d = 0; d < 3; d++) v.velocity[cell * 3 + d] = u_[3 * i + d];
        }
    return v;
  }

  // ---- whole-field access (the device layout is SoA over the lattice) ----
  void getPopulations(std::vector<double>& f) {
    f.resize((size_t)n_ * LB_CURRENT_DIR);
    check(lbm_get_populations(ctx_, f.data()), ctx_, "lbm_get_populations");
  }
  void getFields(std::vector<double>& rho, std::vector<double>& u) {
    rho.resize((size_t)n_);
    u.resize((size_t)n_ * 3);
    check(lbm_get_fields(ctx_, rho.data(), u.data()), ctx_, "lbm_get_fields");
  }
  void reduceMoments(double out[4]) { check(lbm_reduce_moments(ctx_, out), ctx_, "lbm_reduce_moments"); }

  // legacy VTK (structured points) of rho and u per LB cell, the quantities the reference's
  // VTK output mapping plots (BLB/mappings/RegularGrid2BlockVTKOutput.cpp:95-273)
  void writeVTK(const std::string& filename) {
    std::vector<double> rho, u;
    getFields(rho, u);
    FILE* fp = std::fopen(filename.c_str(), "w");
    if (!fp) throw std::runtime_error("cannot open " + filename);
    const int nx = cfg_.numberOfGridPoints[0], ny = cfg_.numberOfGridPoints[1], nz = cfg_.numberOfGridPoints[2];
    const double dx = cfg_.getDx();
    std::fprintf(fp, "# vtk DataFile Version 2.0\nMI355X block-LB fields\nASCII\nDATASET STRUCTURED_POINTS\n");
    std::fprintf(fp, "DIMENSIONS %d %d %d\nORIGIN %.17g %.17g %.17g\nSPACING %.17g %.17g %.17g\n", nx, ny, nz,
                 0.5 * dx, 0.5 * dx, 0.5 * dx, dx, dx, dx);
    std::fprintf(fp, "POINT_DATA %ld\nSCALARS density double 1\nLOOKUP_TABLE default\n", (long)n_);
    for (int64_t i = 0; i < n_; i++) std::fprintf(fp, "%.17g\n", rho[i]);
    std::fprintf(fp, "VECTORS velocity double\n");
    for (int64_t i = 0; i < n_; i++) std::fprintf(fp, "%.17g %.17g %.17g\n", u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    std::fclose(fp);
  }

  lbm_ctx* handle() { return ctx_; }

 private:
  static double lbm_tau_eff(double tau) {  // LB/MultiLevelSimData.h:160
    double nu = 1.0 / 3.0 * (tau - 0.5);
    return 3.0 * nu + 0.5;
  }
  static void check(int rc, const lbm_ctx* c, const char* what) {
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error(c));
  }
  void refreshHostCopy() {
    if (hostStep_ == state_.timestep_ && !f_.empty()) return;
    getPopulations(f_);
    getFields(rho_, u_);
    hostStep_ = state_.timestep_;
  }

  Configuration cfg_;
  lbm_ctx* ctx_ = nullptr;
  int64_t n_ = 0;
  RegularGridBlockState state_;
  bool vtk_ = false;
  std::string vtkBase_;
  double totalSeconds_ = 0.0;
  int64_t iterations_ = 0;
  std::vector<double> f_, rho_, u_;
  int hostStep_ = -1;
};

// The solver mapping's iteration bracket (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:251-340):
// beginIteration resets the cell counter, endIteration runs the fused step and records the counter.
class RegularGrid2RegularBlockSolver {
 public:
  explicit RegularGrid2RegularBlockSolver(Repository& repo) : repo_(repo) {}
  void beginIteration(RegularGridBlockState&) {}
  void endIteration(RegularGridBlockState&) { repo_.iterate(); }

 private:
  Repository& repo_;
};

// BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:49-203
class Runner {
 public:
  int run(const Configuration& cfg) {
    Repository repo(cfg);
    return runAsMaster(repo);
  }
  int runAsMaster(Repository& repo) {
    RegularGridBlockState& st = repo.getRegularGridState();
    for (int t = 0; t < st.getNumberTimesteps(); t++) {
      repo.switchToRegularBlockSolverAdapter();
      repo.iterate();
      // LB/PerformanceMeasurement.cpp:45-66 (MLUPS = inner cells / (1e6 s), wall clock here)
      if (st.measurePerformance()) std::printf("MLUPS: %g\n", st.getLastMLUPS());
    }
    repo.logIterationStatistics();
    return 0;
  }
};

}  // namespace blocklatticeboltzmann
}  // namespace madpac_amd
