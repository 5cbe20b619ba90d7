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
#include <cinttypes>
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
// Storage of the state on the device: two buffers, or one updated in place (half the memory, the
// same results); the reference keeps one block store on the host and has no such choice.
enum class Layout { TwoBuffer = LBM_LAYOUT_TWO_BUFFER, InPlace = LBM_LAYOUT_AA };

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
  Layout layout = Layout::TwoBuffer;
  int device = 0;
  double domainOffset[3] = {0.0, 0.0, 0.0};
  // the plotter tag (LB/configurations/LatticeBoltzmannPlotterConfiguration.cpp:25-55): VTK files
  // <vtkOutputName>_<iteration>.vtk every timestepsPerPlotting iterations; empty name = no plotting
  std::string vtkOutputName;
  int timestepsPerPlotting = 0;

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
  int getTimestep() const { return timestep_; }              // iterations done
  // BLB/RegularGridBlockState.cpp:253-270: with plotting on, returns whether this iteration plots
  // and advances the plot counter
  bool plotVTK() const { return !vtkFilename_.empty(); }
  bool triggerPlotting() {
    if (!plotVTK()) return false;
    const bool plot = plotTimestep_ % timestepsPerPlotting_ == 0;
    plotTimestep_++;
    return plot;
  }
  const std::string& getVTKFilename() const { return vtkFilename_; }
  int64_t getLocalNumberOfCells() const { return cells_; }   // inner cells updated last iteration
  bool measurePerformance() const { return measure_; }
  double getLastMLUPS() const { return mlups_; }

 private:
  friend class Repository;
  double tau_ = 0, viscosityL_ = 0, dx_ = 0, dt_ = 0, density_ = 1, mlups_ = 0;
  int numberTimesteps_ = 0, timestep_ = 0;
  int plotTimestep_ = 0, timestepsPerPlotting_ = 1;
  std::string vtkFilename_;
  int64_t cells_ = 0;
  bool measure_ = false;
};

// Legacy-VTK text file of the per-cell fields, laid out as the reference's block VTK output
// (BLB/mappings/RegularGrid2BlockVTKOutput.cpp:95-190 through
// tarch/plotter/griddata/unstructured/vtk/VTKTextFileWriter.cpp:49-83): an unstructured grid of
// voxels (cell type 11), one per fluid cell, visited block by block in the block store's vertex
// order (6^3 blocks, x fastest; cells z, y, x inside a block); corner points numbered on first
// use, corners of a voxel in x-fastest order; cell data density, velocity and block-id (the
// linear block number, empty blocks counted), values written as float with six significant
// digits.  rho and u are per lattice cell, u interleaved (3 per cell).
inline void writeBlockVTK(const std::string& filename, const int n[3], double dx, const double offset[3],
                          const std::vector<uint8_t>& fluid, const std::vector<double>& rho,
                          const std::vector<double>& u) {
  const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
  if ((int64_t)rho.size() != ncell || (int64_t)u.size() != 3 * ncell || (int64_t)fluid.size() != ncell)
    throw std::runtime_error("writeBlockVTK: field sizes do not match the lattice");
  int nb[3];
  for (int d = 0; d < 3; d++) nb[d] = n[d] / LB_BLOCKSIZE + 1 + (n[d] % LB_BLOCKSIZE ? 1 : 0);
  const int64_t vx = n[0] + 1, vy = n[1] + 1;
  std::vector<int64_t> vnum((size_t)(vx * vy * (n[2] + 1)), -1);
  std::string pts, cells, data_rho, data_u, data_id;
  int64_t npts = 0, ncells = 0;
  char buf[160];
  for (int bz = 0; bz < nb[2]; bz++)
    for (int by = 0; by < nb[1]; by++)
      for (int bx = 0; bx < nb[0]; bx++) {
        const int64_t block = bx + (int64_t)nb[0] * (by + (int64_t)nb[1] * bz);
        for (int z = 0; z < LB_BLOCKSIZE; z++)
          for (int y = 0; y < LB_BLOCKSIZE; y++)
            for (int x = 0; x < LB_BLOCKSIZE; x++) {
              const int g[3] = {LB_BLOCKSIZE * bx + x - 3, LB_BLOCKSIZE * by + y - 3, LB_BLOCKSIZE * bz + z - 3};
              if (g[0] < 0 || g[1] < 0 || g[2] < 0 || g[0] >= n[0] || g[1] >= n[1] || g[2] >= n[2]) continue;
              const int64_t i = g[0] + (int64_t)n[0] * (g[1] + (int64_t)n[1] * g[2]);
              if (!fluid[(size_t)i]) continue;
              int64_t corner[8];
              for (int c = 0; c < 8; c++) {
                const int cx = g[0] + (c & 1), cy = g[1] + ((c >> 1) & 1), cz = g[2] + (c >> 2);
                int64_t& v = vnum[(size_t)(cx + vx * (cy + vy * cz))];
                if (v < 0) {
                  v = npts++;
                  std::snprintf(buf, sizeof buf, "%g  %g  %g\n", offset[0] + cx * dx, offset[1] + cy * dx,
                                offset[2] + cz * dx);
                  pts += buf;
                }
                corner[c] = v;
              }
              std::snprintf(buf, sizeof buf, "8 %" PRId64 " %" PRId64 " %" PRId64 " %" PRId64 " %" PRId64 " %" PRId64
                            " %" PRId64 " %" PRId64 " \n", corner[0], corner[1], corner[2], corner[3], corner[4],
                            corner[5], corner[6], corner[7]);
              cells += buf;
              std::snprintf(buf, sizeof buf, "%g \n", (double)(float)rho[(size_t)i]);
              data_rho += buf;
              std::snprintf(buf, sizeof buf, "%g %g %g \n", (double)(float)u[(size_t)(3 * i)],
                            (double)(float)u[(size_t)(3 * i + 1)], (double)(float)u[(size_t)(3 * i + 2)]);
              data_u += buf;
              std::snprintf(buf, sizeof buf, "%g \n", (double)(float)block);
              data_id += buf;
              ncells++;
            }
      }
  FILE* fp = std::fopen(filename.c_str(), "w");
  if (!fp) throw std::runtime_error("cannot open " + filename);
  std::fprintf(fp, "# vtk DataFile Version 2.0\nMI355X block-LB cell fields\nASCII\n\n");
  std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\nPOINTS %" PRId64 " float\n\n%s\n\n", npts, pts.c_str());
  std::fprintf(fp, "CELLS %" PRId64 " %" PRId64 "\n\n%s\n\n", ncells, 9 * ncells, cells.c_str());
  std::fprintf(fp, "CELL_TYPES %" PRId64 "\n\n", ncells);
  for (int64_t k = 0; k < ncells; k++) std::fputs("11\n", fp);
  std::fprintf(fp, "\n\nCELL_DATA %" PRId64 "\n\n", ncells);
  std::fprintf(fp, "SCALARS density float 1\nLOOKUP_TABLE default\n%s\n", data_rho.c_str());
  std::fprintf(fp, "VECTORS velocity float\n%s\n", data_u.c_str());
  std::fprintf(fp, "SCALARS block-id float 1\nLOOKUP_TABLE default\n%s\n\n\n", data_id.c_str());
  if (std::fclose(fp) != 0) throw std::runtime_error("error writing " + filename);
}

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
    if (!cfg_.vtkOutputName.empty() && cfg_.timestepsPerPlotting < 1)
      throw std::runtime_error("'timesteps-per-plotting' is smaller than one");
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
    c.layout = static_cast<int>(cfg_.layout);
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
    if (!cfg_.vtkOutputName.empty()) {
      state_.vtkFilename_ = cfg_.vtkOutputName;
      state_.timestepsPerPlotting_ = cfg_.timestepsPerPlotting;
    }
    fluid_ = inner;
    // the block store's initial moments (GridManagementService.cpp:121-130)
    heldRho_.assign((size_t)n_, 1.0);
    heldU_.assign((size_t)n_ * 3, 0.0);
  }
  Repository(const Repository&) = delete;
  Repository& operator=(const Repository&) = delete;
  ~Repository() { terminate(); }

  void switchToRegularBlockSolverAdapter() { vtk_ = false; }
  // the file name comes from the state (the configuration's vtkOutputName)
  void switchToRegularBlockSolverAndVTKOutputAdapter() {
    if (!state_.plotVTK()) throw std::runtime_error("no VTK output name configured");
    vtk_ = true;
  }

  // One time step: beginIteration, the fused A/B/C sweeps, endIteration.  With the VTK adapter the
  // file shows what the reference's block store holds when its output mapping runs: the density
  // and velocity its previous iteration computed before colliding (the state one step back; the
  // initial rho = 1, u = 0 in the first iteration), named <vtkOutputName>_<iteration>.vtk.
  void iterate() {
    if (state_.plotVTK()) {
      if (vtk_) {
        const int n3[3] = {cfg_.numberOfGridPoints[0], cfg_.numberOfGridPoints[1], cfg_.numberOfGridPoints[2]};
        writeBlockVTK(state_.getVTKFilename() + "_" + std::to_string(state_.timestep_) + ".vtk", n3, cfg_.getDx(),
                      cfg_.domainOffset, fluid_, heldRho_, heldU_);
      }
      getFields(heldRho_, heldU_);  // the moments this iteration's collision computes
    }
    const auto t0 = std::chrono::steady_clock::now();
    check(lbm_step(ctx_, 1), ctx_, "lbm_step");
    check(lbm_synchronize(ctx_), ctx_, "lbm_synchronize");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    state_.cells_ = lbm_fluid_cells(ctx_);
    state_.mlups_ = s > 0 ? (double)state_.cells_ / (1e6 * s) : 0.0;
    totalSeconds_ += s;
    iterations_++;
    state_.timestep_++;
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

  // the current state's fields in the block VTK layout (writeBlockVTK)
  void writeVTK(const std::string& filename) {
    std::vector<double> rho, u;
    getFields(rho, u);
    const int n3[3] = {cfg_.numberOfGridPoints[0], cfg_.numberOfGridPoints[1], cfg_.numberOfGridPoints[2]};
    writeBlockVTK(filename, n3, cfg_.getDx(), cfg_.domainOffset, fluid_, rho, u);
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
  std::vector<uint8_t> fluid_;
  std::vector<double> heldRho_, heldU_;
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
      if (st.triggerPlotting())
        repo.switchToRegularBlockSolverAndVTKOutputAdapter();
      else
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
