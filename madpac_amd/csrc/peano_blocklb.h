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
//   BLB/mappings/RegularGrid2RegularBlockSolver events beginIteration,
//        touchVertexFirstTime, handleCell, touchVertexLastTime, endIteration          -> RegularGrid2RegularBlockSolver
//   BLB/services/GridManagementService loadVertex/getPdf/getInner/getBoundary/
//        getDensity/getVelocity (6^3 blocks, vertex numbers from 8, written back)     -> GridManagementService
//   peano/geometry/builtin Hexahedron, Sphere, intersected (builtin-geometry tag)     -> Configuration::geometry
//   BLB/BlockRestrictionProlongation (restriction / prolongation between levels)      -> BlockRestrictionProlongation
//   the kernels the mappings call per block / per Peano cell: BLB/BlockComputeCellValues,
//        collisionmodels/BlockBGK, boundaryconditions/BlockBoundaryConditionManager,
//        nonlocaltreatment/AABlockStreaming (batched over vertex lists, device block store)
//        -> DeviceBlocks, BlockComputeCellValues, BlockBGK, BlockBoundaryConditionManager, AABlockStreaming
//   BLB/RegularGridBlockVertex's exchange of a vertex's copies between ranks (Parallel build:
//        init/compute/mergePdfDiff, set/reset/mergeDensityOnBoundary)                -> BlockVertexExchange
//   LB/PerformanceMeasurement (MLUPS per iteration; wall clock here, not CPU time)
// The three sweeps of one reference iteration (per-block collide + local streaming, the
// cross-block streaming of handleCell, the in-cell swap) are one fused device step here.  A mapping
// that wants the per-block and per-cell events gets them around that step: touchVertexFirstTime and
// handleCell see the blocks before it (the state the reference's touchVertexFirstTime loads), and
// touchVertexLastTime after it (the state after the reference's in-cell swap).  Errors that the
// reference reports through assertions throw std::runtime_error.
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include <map>

#include "lbm_capi.h"
#include "peano_geometry.h"

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
  // the builtin-geometry tag (peano/geometry/builtin/configurations/BuiltinGeometryConfiguration.cpp:
  // 25-77): the fluid domain is the intersection of these objects; an inverted object is an
  // obstacle.  Empty means one hexahedron covering the domain box, the cavity and channel scenarios'
  // usual geometry.  Cells whose centre lies outside the closed domain are solid (halfway bounce-back).
  std::vector<geometry::Object> geometry;

  geometry::Domain domain() const {
    geometry::Domain d;
    if (geometry.empty()) {
      d.objects.push_back(geometry::Object::hexahedron({domainOffset[0], domainOffset[1], domainOffset[2]},
                                                       {domainSize[0], domainSize[1], domainSize[2]}));
    } else {
      d.objects = geometry;
    }
    return d;
  }

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

// A vertex of the regular grid, i.e. one block: its number in the block store and its position
// (BLB/RegularGridBlockVertex::getVertexNumber; the vertex position the grid passes to the events).
class RegularGridBlockVertex {
 public:
  unsigned getVertexNumber() const { return number_; }
  const geometry::Vec3& getX() const { return x_; }

 private:
  friend class Repository;
  unsigned number_ = 0;
  geometry::Vec3 x_{};
};
// A cell of the regular grid between 2^3 vertices; the block solver keeps no data on it.
class RegularGridBlockCell {};
// The cell's vertices are vertices[e(i)] for i = 0..7 in lexicographic order (x fastest), as
// peano/kernel/gridinterface/VertexEnumerator orders them.
class VertexEnumerator {
 public:
  int operator()(int i) const { return i; }
  geometry::Vec3 getVertexPosition(int i = 0) const {
    geometry::Vec3 x = x0_;
    for (int d = 0; d < 3; d++) x[d] += (double)((i >> d) & 1) * h_[d];
    return x;
  }
  geometry::Vec3 getCellSize() const { return h_; }

 private:
  friend class Repository;
  geometry::Vec3 x0_{}, h_{};
};

// The solver mapping: the events of one iteration (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:
// touchVertexFirstTime :83-177, handleCell :180-220, touchVertexLastTime :223-248, beginIteration
// :251-310, endIteration :313-340).  The reference's own work in those events (collide + local
// streaming, cross-block streaming, in-cell swap) is the fused device step, run by the repository
// between touchVertexFirstTime/handleCell and touchVertexLastTime; this base class adds nothing
// to it.  A derived mapping that overrides the per-block or per-cell events returns true from
// hasBlockEvents(): the repository then visits every vertex and cell on the host each iteration,
// with block data through GridManagementService (a host copy of the state, written back before the
// step).  Without it (the default) an iteration is only the device step.
class RegularGrid2RegularBlockSolver {
 public:
  virtual ~RegularGrid2RegularBlockSolver() = default;
  virtual bool hasBlockEvents() const { return false; }
  virtual void beginIteration(RegularGridBlockState&) {}
  virtual void touchVertexFirstTime(RegularGridBlockVertex&, const geometry::Vec3&) {}
  virtual void handleCell(RegularGridBlockVertex* const, RegularGridBlockCell&, const VertexEnumerator&) {}
  virtual void touchVertexLastTime(RegularGridBlockVertex&, const geometry::Vec3&) {}
  virtual void endIteration(RegularGridBlockState&) {}
};

// Restriction and prolongation between grid levels (BLB/BlockRestrictionProlongation.h:34-155), with
// the reference's method names and block arguments (6^3 blocks in GridManagementService's layout,
// std::vector in place of tarch::la::Vector, child positions 0..3 per axis).  Each call runs on the
// device through an lbm_blockstore of two blocks (a host round trip per call, as the reference's
// per-vertex calls are); a spacetree traversal that collects the vertices of one level pair calls
// lbm_blocks_find_overlap / _prolongate / _restrict once for all of them instead.
class BlockRestrictionProlongation {
 public:
  explicit BlockRestrictionProlongation(int device = 0) {
    if (lbm_blocks_create(&bs_, 2, device) != LBM_OK)
      throw std::runtime_error(std::string("lbm_blocks_create: ") + lbm_blocks_last_error(nullptr));
  }
  ~BlockRestrictionProlongation() { lbm_blocks_destroy(bs_); }
  BlockRestrictionProlongation(const BlockRestrictionProlongation&) = delete;
  BlockRestrictionProlongation& operator=(const BlockRestrictionProlongation&) = delete;

  // :17-28
  std::array<int, 3> getFatherVertex(const std::array<int, 3>& discretePosition) const {
    return {discretePosition[0] / 2, discretePosition[1] / 2, discretePosition[2] / 2};
  }
  // :264-288
  std::array<int, 3> getSubBlock(const std::array<int, 3>& discretePosition) const {
    static const int t[4] = {1, 2, 0, 1};
    return {t[discretePosition[0]], t[discretePosition[1]], t[discretePosition[2]]};
  }
  // :32-62; hasMinusOneEntries gets 216 flags
  void findFineOverlapCells(const std::vector<double>& finePdf, std::vector<bool>& hasMinusOneEntries) const {
    size(finePdf, kPdf, "finePdf");
    check(lbm_blocks_set(bs_, 1, 1, finePdf.data(), nullptr, nullptr), "lbm_blocks_set");
    const lbm_amr_link l = link({0, 0, 0});
    uint8_t m[LB_BLOCK_NUMBER_OF_CELLS];
    check(lbm_blocks_find_overlap(bs_, &l, 1, m), "lbm_blocks_find_overlap");
    hasMinusOneEntries.assign(m, m + LB_BLOCK_NUMBER_OF_CELLS);
  }
  // :83-160
  void prolongateToFineLevel(const std::array<int, 3>& discretePosition, const std::vector<double>& coarsePdf,
                             std::vector<double>& finePdf, std::vector<double>& density, std::vector<double>& velocity,
                             const std::vector<bool>& hasMinusOneEntries) const {
    size(coarsePdf, kPdf, "coarsePdf");
    size(finePdf, kPdf, "finePdf");
    size(density, LB_BLOCK_NUMBER_OF_CELLS, "density");
    size(velocity, 3 * LB_BLOCK_NUMBER_OF_CELLS, "velocity");
    if (hasMinusOneEntries.size() != (size_t)LB_BLOCK_NUMBER_OF_CELLS)
      throw std::runtime_error("hasMinusOneEntries must hold 216 flags");
    uint8_t m[LB_BLOCK_NUMBER_OF_CELLS];
    for (int c = 0; c < LB_BLOCK_NUMBER_OF_CELLS; c++) m[c] = hasMinusOneEntries[c] ? 1 : 0;
    check(lbm_blocks_set(bs_, 0, 1, coarsePdf.data(), nullptr, nullptr), "lbm_blocks_set");
    check(lbm_blocks_set(bs_, 1, 1, nullptr, density.data(), velocity.data()), "lbm_blocks_set");
    check(lbm_blocks_set_overlap(bs_, 1, 1, m), "lbm_blocks_set_overlap");
    const lbm_amr_link l = link(discretePosition);
    check(lbm_blocks_set(bs_, 1, 1, finePdf.data(), nullptr, nullptr), "lbm_blocks_set");
    check(lbm_blocks_prolongate(bs_, &l, 1), "lbm_blocks_prolongate");
    check(lbm_blocks_get(bs_, 1, 1, finePdf.data(), density.data(), velocity.data()), "lbm_blocks_get");
  }
  // :164-259
  void restrictToCoarseLevel(const std::array<int, 3>& discretePosition, std::vector<double>& coarsePdf,
                             const std::vector<double>& finePdf) const {
    size(coarsePdf, kPdf, "coarsePdf");
    size(finePdf, kPdf, "finePdf");
    check(lbm_blocks_set(bs_, 0, 1, coarsePdf.data(), nullptr, nullptr), "lbm_blocks_set");
    check(lbm_blocks_set(bs_, 1, 1, finePdf.data(), nullptr, nullptr), "lbm_blocks_set");
    const lbm_amr_link l = link(discretePosition);
    check(lbm_blocks_restrict(bs_, &l, 1), "lbm_blocks_restrict");
    check(lbm_blocks_get(bs_, 0, 1, coarsePdf.data(), nullptr, nullptr), "lbm_blocks_get");
  }

 private:
  static constexpr int kPdf = LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR;
  static lbm_amr_link link(const std::array<int, 3>& pos) {
    lbm_amr_link l{};
    l.coarse = 0;
    l.fine = 1;
    for (int d = 0; d < 3; d++) l.pos[d] = pos[d];
    return l;
  }
  static void size(const std::vector<double>& v, int n, const char* what) {
    if (v.size() != (size_t)n) throw std::runtime_error(std::string(what) + ": wrong size");
  }
  void check(int rc, const char* what) const {
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_blocks_last_error(bs_));
  }
  lbm_blockstore* bs_ = nullptr;
};

// The block solver's kernels for code that drives them itself, as faxen/lbf's solver mapping reuses
// them (peano/applications/faxen/lbf/mappings/RegularGrid2RegularBlockSolver.cpp:111,184,205): the
// reference's classes and method names, with the blocks in a device store addressed by vertex
// number instead of one host block per call.  Each method takes the list of vertices (or Peano
// cells) that the caller's traversal would visit one at a time and runs them in one launch, so a
// mapping collects its touchVertexFirstTime vertices, its handleCell cells and its
// touchVertexLastTime vertices and makes one call per kernel.  A vertex appears once per list and
// at one vertex position once per cell list (lbm_blocks_* return LBM_ERR_ARG otherwise, thrown
// here as std::runtime_error).
class DeviceBlocks {
 public:
  explicit DeviceBlocks(int64_t nblocks, int device = 0) {
    if (lbm_blocks_create(&bs_, nblocks, device) != LBM_OK)
      throw std::runtime_error(std::string("lbm_blocks_create: ") + lbm_blocks_last_error(nullptr));
  }
  ~DeviceBlocks() { lbm_blocks_destroy(bs_); }
  DeviceBlocks(const DeviceBlocks&) = delete;
  DeviceBlocks& operator=(const DeviceBlocks&) = delete;
  lbm_blockstore* handle() const { return bs_; }
  // one vertex's block in or out (GridManagementService::VertexData: pdf, density, velocity, inner)
  void setBlock(int64_t vertex, const BlockView& b) {
    check(lbm_blocks_set(bs_, vertex, 1, b.pdf.empty() ? nullptr : b.pdf.data(),
                         b.density.empty() ? nullptr : b.density.data(), b.velocity.empty() ? nullptr : b.velocity.data()),
          "lbm_blocks_set");
    if (!b.inner.empty()) check(lbm_blocks_set_inner(bs_, vertex, 1, b.inner.data()), "lbm_blocks_set_inner");
  }
  void getBlock(int64_t vertex, BlockView& b) const {
    b.vertexNumber = (unsigned)vertex;
    b.pdf.resize(LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR);
    b.density.resize(LB_BLOCK_NUMBER_OF_CELLS);
    b.velocity.resize(3 * LB_BLOCK_NUMBER_OF_CELLS);
    check(lbm_blocks_get(bs_, vertex, 1, b.pdf.data(), b.density.data(), b.velocity.data()), "lbm_blocks_get");
  }
  void check(int rc, const char* what) const {
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_blocks_last_error(bs_));
  }

 private:
  lbm_blockstore* bs_ = nullptr;
};

// BLB/BlockComputeCellValues.cpp:26-68: computeDensity then computeVelocity on each listed block;
// hasMinusOneEntries = the blocks' overlap cells of their last overlap search
class BlockComputeCellValues {
 public:
  void computeDensityAndVelocity(DeviceBlocks& s, const std::vector<int64_t>& vertices,
                                 bool hasMinusOneEntries = false) const {
    s.check(lbm_blocks_compute_cell_values(s.handle(), vertices.data(), (int64_t)vertices.size(), hasMinusOneEntries),
            "lbm_blocks_compute_cell_values");
  }
};

// BLB/collisionmodels/BlockBGK.cpp:18-45 through CollisionModelWrapper::collide (:28-39)
class BlockBGK {
 public:
  explicit BlockBGK(double tau) : tau_(tau) {}
  double getTau() const { return tau_; }
  void collide(DeviceBlocks& s, const std::vector<int64_t>& vertices, bool hasMinusOneEntries = false) const {
    s.check(lbm_blocks_collide(s.handle(), vertices.data(), (int64_t)vertices.size(), tau_, hasMinusOneEntries),
            "lbm_blocks_collide");
  }

 private:
  double tau_;
};

// BLB/boundaryconditions/BlockBoundaryConditionManager.cpp:28-62; the links are the
// (BoundaryData, boundaryDataIndex) pairs of ReceiveBoundaryDataService::getBoundaryData with the
// block they belong to
class BlockBoundaryConditionManager {
 public:
  void treatBoundaryBeforeCollision(DeviceBlocks&, const std::vector<lbm_block_boundary>&) const {}  // no-op for every condition
  void treatBoundaryAfterCollision(DeviceBlocks& s, const std::vector<lbm_block_boundary>& links) const {
    s.check(lbm_blocks_treat_boundary(s.handle(), links.data(), (int64_t)links.size()), "lbm_blocks_treat_boundary");
  }
};

// BLB/nonlocaltreatment/AABlockStreaming.cpp:19-92; a Peano cell is its 8 vertices in lexicographic
// order (vertices[enumerator(i)], i = 0..7)
class AABlockStreaming {
 public:
  void streamingLocal(DeviceBlocks& s, const std::vector<int64_t>& vertices) const {
    s.check(lbm_blocks_stream_local(s.handle(), vertices.data(), (int64_t)vertices.size()), "lbm_blocks_stream_local");
  }
  void streamingNonLocal(DeviceBlocks& s, const std::vector<std::array<int64_t, 8>>& cells) const {
    s.check(lbm_blocks_stream_nonlocal(s.handle(), cells.empty() ? nullptr : cells[0].data(), (int64_t)cells.size()),
            "lbm_blocks_stream_nonlocal");
  }
  void swapPdfsInsideCells(DeviceBlocks& s, const std::vector<int64_t>& vertices) const {
    s.check(lbm_blocks_swap_inside(s.handle(), vertices.data(), (int64_t)vertices.size()), "lbm_blocks_swap_inside");
  }
};

// The Parallel build's exchange of a vertex's copies between ranks (BLB/RegularGridBlockVertex.cpp:
// 51-160, called from RegularGrid2RegularBlockSolver.cpp:158-159, :243-245 and mergeWithNeighbour
// :349-364): the vertex methods, batched over the vertices a rank shares.  The differences
// (LB_PDFS_ON_BLOCKBOUNDARY = 1,008 per vertex) and boundary densities (208 per vertex) stay in the
// device store between calls; getPdfDiff / getDensityOnBoundary copy them, in list order, into the
// buffer a transport sends, and the merges take the neighbour rank's buffer (host or device memory).
class BlockVertexExchange {
 public:
  static constexpr int PdfsOnBlockBoundary = 1008;
  static constexpr int CellsOnBlockBoundary = 208;
  // the geometry-boundary bitset of one vertex (GridManagementService::getBoundary)
  void setBoundary(DeviceBlocks& s, int64_t vertex, const std::vector<uint8_t>& boundary) const {
    if (boundary.size() != (size_t)LB_BLOCK_NUMBER_OF_CELLS)
      throw std::runtime_error("setBoundary: one flag per cell of the block (216) expected");
    s.check(lbm_blocks_set_boundary(s.handle(), vertex, 1, boundary.data()), "lbm_blocks_set_boundary");
  }
  void initPdfDiff(DeviceBlocks& s, const std::vector<int64_t>& vertices) const {
    s.check(lbm_blocks_init_pdf_diff(s.handle(), vertices.data(), (int64_t)vertices.size()), "lbm_blocks_init_pdf_diff");
  }
  void computePdfDiff(DeviceBlocks& s, const std::vector<int64_t>& vertices) const {
    s.check(lbm_blocks_compute_pdf_diff(s.handle(), vertices.data(), (int64_t)vertices.size()),
            "lbm_blocks_compute_pdf_diff");
  }
  void getPdfDiff(DeviceBlocks& s, const std::vector<int64_t>& vertices, double* out) const {
    s.check(lbm_blocks_get_pdf_diff(s.handle(), vertices.data(), (int64_t)vertices.size(), out), "lbm_blocks_get_pdf_diff");
  }
  void mergePdfDiff(DeviceBlocks& s, const std::vector<int64_t>& vertices, const double* neighbour) const {
    s.check(lbm_blocks_merge_pdf_diff(s.handle(), vertices.data(), (int64_t)vertices.size(), neighbour),
            "lbm_blocks_merge_pdf_diff");
  }
  void resetDensityOnBlockBoundary(DeviceBlocks& s, const std::vector<int64_t>& vertices) const {
    s.check(lbm_blocks_reset_density_on_boundary(s.handle(), vertices.data(), (int64_t)vertices.size()),
            "lbm_blocks_reset_density_on_boundary");
  }
  void setDensityOnBlockBoundary(DeviceBlocks& s, const std::vector<int64_t>& vertices, bool useDynamicGeometry) const {
    s.check(lbm_blocks_set_density_on_boundary(s.handle(), vertices.data(), (int64_t)vertices.size(), useDynamicGeometry),
            "lbm_blocks_set_density_on_boundary");
  }
  void getDensityOnBoundary(DeviceBlocks& s, const std::vector<int64_t>& vertices, double* out) const {
    s.check(lbm_blocks_get_density_on_boundary(s.handle(), vertices.data(), (int64_t)vertices.size(), out),
            "lbm_blocks_get_density_on_boundary");
  }
  void mergeDensityOnBoundary(DeviceBlocks& s, const std::vector<int64_t>& vertices, const double* neighbour) const {
    s.check(lbm_blocks_merge_density_on_boundary(s.handle(), vertices.data(), (int64_t)vertices.size(), neighbour),
            "lbm_blocks_merge_density_on_boundary");
  }
};

class Repository;

// The block store (BLB/services/GridManagementService.h:40-86, .cpp:295-327): loadVertex selects a
// block, the getters return references to its data that callers may write, as faxen/lbf does
// (peano/applications/faxen/lbf/mappings/RegularGrid2RegularBlockSolver.cpp:103,162,229).  Writes to
// pdf and inner reach the device before the next step or read of the state; inner flags that
// change re-flag the lattice (a cell that becomes inner takes the populations written for it, 0
// unless written).  density and velocity are outputs, recomputed from pdf at every step as the
// reference's touchVertexFirstTime does (computeDensity/computeVelocity); writes to them and to
// boundary are kept in the host copy only.  One instance per repository; not thread-safe.
class GridManagementService {
 public:
  explicit GridManagementService(Repository& repo) : repo_(repo) {}
  void loadVertex(unsigned vertexNumber);
  unsigned getVertexNumber() const { return vertex_; }
  std::vector<double>& getPdf();
  std::vector<uint8_t>& getInner();
  std::vector<uint8_t>& getBoundary();
  std::vector<double>& getDensity();
  std::vector<double>& getVelocity();

 private:
  BlockView& current();
  Repository& repo_;
  unsigned vertex_ = 0;
};

class Repository {
 public:
  explicit Repository(const Configuration& cfg) : cfg_(cfg), store_(*this) {
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
    // the block store's flags from the configured geometry (GridManagementService::createVertex,
    // BLB/services/GridManagementService.cpp:86-118); lattice cell c is block slot c + 3
    const geometry::BlockGrid grid(cfg_.numberOfGridPoints, cfg_.getDx(),
                                   {cfg_.domainOffset[0], cfg_.domainOffset[1], cfg_.domainOffset[2]});
    grid_h_ = grid.h;
    gridOffset_ = grid.offset;
    const geometry::Domain dom = cfg_.domain();
    grid.storeFlags(dom, storeInner_, storeBoundary_);
    fluid_ = grid.latticeInner(storeInner_);
    check(lbm_create(&ctx_, &c), nullptr, "lbm_create");
    n_ = (int64_t)c.nx * c.ny * c.nz;
    check(lbm_set_flags(ctx_, fluid_.data()), ctx_, "lbm_set_flags");
    if (cfg_.scenario == Scenario::LidDrivenCavity)
      check(lbm_set_cavity_lid(ctx_, cfg_.characteristicVelocityL), ctx_, "lbm_set_cavity_lid");
    else if (cfg_.scenario == Scenario::DefaultChannel)
      check(lbm_set_channel(ctx_, cfg_.characteristicVelocityL, cfg_.useParabolicProfile ? 1 : 0), ctx_,
            "lbm_set_channel");
    // f = w, rho = 1, u = 0 in inner cells (GridManagementService.cpp:121-148)
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
    // the block store's initial moments (GridManagementService.cpp:121-130)
    heldRho_.assign((size_t)n_, 0.0);
    heldU_.assign((size_t)n_ * 3, 0.0);
    for (int64_t i = 0; i < n_; i++) heldRho_[(size_t)i] = fluid_[(size_t)i] ? 1.0 : 0.0;
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
  // the mapping whose events iterate() fires (nullptr: the plain solver mapping); not owned
  void setMapping(RegularGrid2RegularBlockSolver* mapping) { mapping_ = mapping ? mapping : &plain_; }

  // One iteration: beginIteration; with block events, touchVertexFirstTime for every vertex and
  // handleCell for every cell on the state before the step; the fused A/B/C step; with block
  // events, touchVertexLastTime for every vertex on the state after it; endIteration.  With the
  // VTK adapter the file shows what the reference's block store holds when its output mapping runs:
  // the density and velocity its previous iteration computed before colliding (the state one step
  // back; the initial rho = 1, u = 0 in the first iteration), named <vtkOutputName>_<iteration>.vtk.
  void iterate() {
    RegularGrid2RegularBlockSolver& m = *mapping_;
    sync();
    m.beginIteration(state_);
    const bool events = m.hasBlockEvents();
    if (events) {
      visitVertices(m, true);
      visitCells(m);
      sync();
    }
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
    if (events) visitVertices(m, false);  // writes reach the device at the next sync
    m.endIteration(state_);
  }

  RegularGridBlockState& getRegularGridState() { return state_; }
  GridManagementService& getGridManagementService() { return store_; }

  void terminate() {
    if (ctx_) lbm_destroy(ctx_);
    ctx_ = nullptr;
  }

  void logIterationStatistics() const {
    std::printf("|| adapter name || iterations || total wall time [s] || average wall time [s]\n");
    std::printf("| RegularBlockSolverAdapter | %ld | %g | %g\n", (long)iterations_, totalSeconds_,
                iterations_ ? totalSeconds_ / iterations_ : 0.0);
  }

  // ---- block store (GridManagementService) ----
  // The reference places vertices one Peano cell (6 dx) apart starting at the domain corner and
  // gives each a 6^3 block centred on it, blocks(d) = N/6 + 1 (+1 if N%6) per dimension
  // (BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:78-92); block slot s holds the
  // cell centred at (s - 2.5) dx, i.e. lattice cell s - 3.  Vertex numbers start at 8 (0..7 are the
  // temporary vertices, GridManagementService.cpp:26-54) in lexicographic block order.
  int blocksPerDimension(int d) const {
    const int n = cfg_.numberOfGridPoints[d];
    return n / LB_BLOCKSIZE + 1 + (n % LB_BLOCKSIZE ? 1 : 0);
  }
  unsigned numberOfVertices() const {
    return (unsigned)(blocksPerDimension(0) * blocksPerDimension(1) * blocksPerDimension(2));
  }
  // a copy of block `vertexNumber` of the current state
  BlockView loadVertex(unsigned vertexNumber) { return block(vertexNumber); }

  // The block held for GridManagementService: built from a host copy of the current state on first
  // use, kept (with the caller's writes) until the next sync.
  BlockView& block(unsigned vertexNumber) {
    if (vertexNumber < 8 || vertexNumber >= 8 + numberOfVertices())
      throw std::runtime_error("Could not find vertex " + std::to_string(vertexNumber));
    auto it = blocks_.find(vertexNumber);
    if (it != blocks_.end()) return it->second;
    refreshHostCopy();
    const int b0 = blocksPerDimension(0), b1 = blocksPerDimension(1);
    const unsigned k = vertexNumber - 8;
    const int bx = (int)(k % b0), by = (int)((k / b0) % b1), bz = (int)(k / (b0 * b1));
    BlockView& v = blocks_[vertexNumber];
    v.vertexNumber = vertexNumber;
    v.inner.assign(LB_BLOCK_NUMBER_OF_CELLS, 0);
    v.boundary.assign(LB_BLOCK_NUMBER_OF_CELLS, 0);
    v.pdf.assign(LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, 0.0);
    v.density.assign(LB_BLOCK_NUMBER_OF_CELLS, 0.0);
    v.velocity.assign(LB_BLOCK_NUMBER_OF_CELLS * 3, 0.0);
    for (int cell = 0; cell < LB_BLOCK_NUMBER_OF_CELLS; cell++) {
      const int64_t slot = storeSlot(bx, by, bz, cell);
      v.inner[cell] = storeInner_[(size_t)slot];
      v.boundary[cell] = storeBoundary_[(size_t)slot];
      const int64_t i = latticeCell(bx, by, bz, cell);
      if (i < 0 || !fluid_[(size_t)i]) continue;
      for (int q = 0; q < LB_CURRENT_DIR; q++) v.pdf[cell * LB_CURRENT_DIR + q] = f_[(size_t)q * n_ + i];
      v.density[cell] = rho_[(size_t)i];
      for (int d = 0; d < 3; d++) v.velocity[cell * 3 + d] = u_[(size_t)(3 * i + d)];
    }
    return v;
  }
  void markDirty() { dirty_ = true; }

  // Write the blocks held for GridManagementService back to the device: pdf of lattice cells, and
  // inner flags (re-flagging the lattice when they changed).  Then the held blocks are dropped.
  void sync() {
    if (!dirty_) {
      blocks_.clear();
      return;
    }
    std::vector<uint8_t> inner = fluid_;
    const int b0 = blocksPerDimension(0), b1 = blocksPerDimension(1);
    for (auto& kv : blocks_) {
      const unsigned k = kv.first - 8;
      const int bx = (int)(k % b0), by = (int)((k / b0) % b1), bz = (int)(k / (b0 * b1));
      const BlockView& v = kv.second;
      for (int cell = 0; cell < LB_BLOCK_NUMBER_OF_CELLS; cell++) {
        const int64_t slot = storeSlot(bx, by, bz, cell);
        storeBoundary_[(size_t)slot] = v.boundary[cell] ? 1 : 0;
        const int64_t i = latticeCell(bx, by, bz, cell);
        if (i < 0) {
          if (v.inner[cell]) throw std::runtime_error("a block-store cell outside the lattice cannot be inner");
          continue;
        }
        storeInner_[(size_t)slot] = v.inner[cell] ? 1 : 0;
        inner[(size_t)i] = v.inner[cell] ? 1 : 0;
        for (int q = 0; q < LB_CURRENT_DIR; q++)
          f_[(size_t)q * n_ + i] = inner[(size_t)i] ? v.pdf[cell * LB_CURRENT_DIR + q] : 0.0;
      }
    }
    blocks_.clear();
    dirty_ = false;
    if (inner != fluid_) {
      fluid_ = inner;
      check(lbm_set_flags(ctx_, fluid_.data()), ctx_, "lbm_set_flags");
    }
    check(lbm_set_populations(ctx_, f_.data()), ctx_, "lbm_set_populations");
    hostStep_ = -1;  // the device may round what it stores (FP32-shifted contexts)
  }

  // ---- whole-field access (the device layout is SoA over the lattice) ----
  void getPopulations(std::vector<double>& f) {
    sync();
    f.resize((size_t)n_ * LB_CURRENT_DIR);
    check(lbm_get_populations(ctx_, f.data()), ctx_, "lbm_get_populations");
  }
  void getFields(std::vector<double>& rho, std::vector<double>& u) {
    sync();
    rho.resize((size_t)n_);
    u.resize((size_t)n_ * 3);
    check(lbm_get_fields(ctx_, rho.data(), u.data()), ctx_, "lbm_get_fields");
  }
  void reduceMoments(double out[4]) {
    sync();
    check(lbm_reduce_moments(ctx_, out), ctx_, "lbm_reduce_moments");
  }
  // the inner flags of the lattice cells (row-major)
  const std::vector<uint8_t>& fluidFlags() const { return fluid_; }

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
    f_.resize((size_t)n_ * LB_CURRENT_DIR);
    rho_.resize((size_t)n_);
    u_.resize((size_t)n_ * 3);
    check(lbm_get_populations(ctx_, f_.data()), ctx_, "lbm_get_populations");
    check(lbm_get_fields(ctx_, rho_.data(), u_.data()), ctx_, "lbm_get_fields");
    hostStep_ = state_.timestep_;
  }
  // block-store slot of (block, cell) in the [6 b2][6 b1][6 b0] flag arrays
  int64_t storeSlot(int bx, int by, int bz, int cell) const {
    const int x = cell % 6, y = (cell / 6) % 6, z = cell / 36;
    const int64_t sx = 6 * blocksPerDimension(0), sy = 6 * blocksPerDimension(1);
    return (6 * bx + x) + sx * ((6 * by + y) + sy * (int64_t)(6 * bz + z));
  }
  // lattice cell of (block, cell), or -1 outside the lattice
  int64_t latticeCell(int bx, int by, int bz, int cell) const {
    const int gx = 6 * bx + cell % 6 - 3, gy = 6 * by + (cell / 6) % 6 - 3, gz = 6 * bz + cell / 36 - 3;
    const int nx = cfg_.numberOfGridPoints[0], ny = cfg_.numberOfGridPoints[1], nz = cfg_.numberOfGridPoints[2];
    if (gx < 0 || gy < 0 || gz < 0 || gx >= nx || gy >= ny || gz >= nz) return -1;
    return gx + (int64_t)nx * (gy + (int64_t)ny * gz);
  }
  RegularGridBlockVertex vertexAt(int bx, int by, int bz) const {
    RegularGridBlockVertex v;
    v.number_ = 8u + (unsigned)(bx + blocksPerDimension(0) * (by + blocksPerDimension(1) * bz));
    const int b[3] = {bx, by, bz};
    v.x_ = gridOffset_;
    for (int d = 0; d < 3; d++) v.x_[d] += (double)b[d] * grid_h_[d];
    return v;
  }
  // the per-vertex events in the grid's lexicographic order (first or last touch)
  void visitVertices(RegularGrid2RegularBlockSolver& m, bool first) {
    for (int bz = 0; bz < blocksPerDimension(2); bz++)
      for (int by = 0; by < blocksPerDimension(1); by++)
        for (int bx = 0; bx < blocksPerDimension(0); bx++) {
          RegularGridBlockVertex v = vertexAt(bx, by, bz);
          const geometry::Vec3 x = v.getX();
          if (first) m.touchVertexFirstTime(v, x);
          else m.touchVertexLastTime(v, x);
        }
  }
  // handleCell for every cell between 2^3 vertices
  void visitCells(RegularGrid2RegularBlockSolver& m) {
    RegularGridBlockCell cell;
    for (int cz = 0; cz + 1 < blocksPerDimension(2); cz++)
      for (int cy = 0; cy + 1 < blocksPerDimension(1); cy++)
        for (int cx = 0; cx + 1 < blocksPerDimension(0); cx++) {
          RegularGridBlockVertex vs[8];
          for (int i = 0; i < 8; i++) vs[i] = vertexAt(cx + (i & 1), cy + ((i >> 1) & 1), cz + (i >> 2));
          VertexEnumerator e;
          e.x0_ = vs[0].getX();
          e.h_ = grid_h_;
          m.handleCell(vs, cell, e);
        }
  }

  Configuration cfg_;
  lbm_ctx* ctx_ = nullptr;
  int64_t n_ = 0;
  RegularGridBlockState state_;
  bool vtk_ = false;
  std::vector<uint8_t> fluid_;                       // lattice inner flags (row-major)
  std::vector<uint8_t> storeInner_, storeBoundary_;  // every block-store slot
  geometry::Vec3 grid_h_{}, gridOffset_{};
  std::vector<double> heldRho_, heldU_;
  double totalSeconds_ = 0.0;
  int64_t iterations_ = 0;
  std::vector<double> f_, rho_, u_;
  int hostStep_ = -1;
  std::map<unsigned, BlockView> blocks_;
  bool dirty_ = false;
  GridManagementService store_;
  RegularGrid2RegularBlockSolver plain_;
  RegularGrid2RegularBlockSolver* mapping_ = &plain_;
};

inline void GridManagementService::loadVertex(unsigned vertexNumber) {
  repo_.block(vertexNumber);  // throws "Could not find vertex" like the reference's store
  vertex_ = vertexNumber;
}
inline BlockView& GridManagementService::current() {
  if (vertex_ < 8) throw std::runtime_error("no vertex loaded");
  return repo_.block(vertex_);
}
inline std::vector<double>& GridManagementService::getPdf() {
  BlockView& v = current();
  repo_.markDirty();
  return v.pdf;
}
inline std::vector<uint8_t>& GridManagementService::getInner() {
  BlockView& v = current();
  repo_.markDirty();
  return v.inner;
}
inline std::vector<uint8_t>& GridManagementService::getBoundary() {
  BlockView& v = current();
  repo_.markDirty();
  return v.boundary;
}
inline std::vector<double>& GridManagementService::getDensity() { return current().density; }
inline std::vector<double>& GridManagementService::getVelocity() { return current().velocity; }

// BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:49-203
class Runner {
 public:
  // mapping: the solver mapping whose events each iteration fires (nullptr: the plain one)
  int run(const Configuration& cfg, RegularGrid2RegularBlockSolver* mapping = nullptr) {
    Repository repo(cfg);
    repo.setMapping(mapping);
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
