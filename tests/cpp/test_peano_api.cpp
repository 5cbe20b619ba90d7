// Drives the reference-shaped C++ API (madpac_amd/csrc/peano_blocklb.h) like reference code drives
// the Peano block-LB: configure, run the time loop, read blocks through the block-store view.
// Usage: test_peano_api cavity|channel|channel-parabolic nx ny nz tau u steps out.bin [vtkname every]
// (writes the SoA populations, FP64; with vtkname, VTK field files every `every` iterations as the
// reference's plotter tag does).  Exit code 0 = the block view agrees with the dense fields and the
// flag layout of the reference.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "peano_blocklb.h"

using namespace madpac_amd::blocklatticeboltzmann;

int main(int argc, char** argv) {
  if (argc != 9 && argc != 11) {
    std::fprintf(stderr, "usage: %s cavity|channel|channel-parabolic nx ny nz tau u steps out.bin [vtkname every]\n",
                 argv[0]);
    return 2;
  }
  Configuration cfg;
  const std::string scenario = argv[1];
  int n3[3];
  for (int d = 0; d < 3; d++) cfg.numberOfGridPoints[d] = n3[d] = std::atoi(argv[2 + d]);
  for (int d = 0; d < 3; d++) cfg.domainSize[d] = (double)n3[d] / n3[1];
  cfg.tau = std::atof(argv[5]);
  cfg.characteristicVelocityL = std::atof(argv[6]);
  cfg.numberOfTimesteps = std::atoi(argv[7]);
  const char* outpath = argv[8];
  if (argc == 11) {
    cfg.vtkOutputName = argv[9];
    cfg.timestepsPerPlotting = std::atoi(argv[10]);
  }
  // LBM_TEST_LAYOUT=in-place runs the same case with the state updated in place
  if (const char* l = std::getenv("LBM_TEST_LAYOUT"))
    if (std::string(l) == "in-place") cfg.layout = Layout::InPlace;
  if (scenario == "cavity") {
    cfg.scenario = Scenario::LidDrivenCavity;
  } else if (scenario == "channel" || scenario == "channel-parabolic") {
    cfg.scenario = Scenario::DefaultChannel;
    cfg.useParabolicProfile = scenario == "channel-parabolic";
  } else {
    return 2;
  }
  const int64_t ncell = (int64_t)n3[0] * n3[1] * n3[2];
  try {
    Repository repo(cfg);
    Runner runner;
    runner.runAsMaster(repo);
    RegularGridBlockState& st = repo.getRegularGridState();
    if (st.getTimestep() != cfg.numberOfTimesteps) return 3;
    if (st.getLocalNumberOfCells() != ncell) return 4;
    std::vector<double> f, rho, u;
    repo.getPopulations(f);
    repo.getFields(rho, u);
    FILE* fp = std::fopen(outpath, "wb");
    if (!fp) return 5;
    std::fwrite(f.data(), sizeof(double), f.size(), fp);
    std::fclose(fp);
    // block view: every block, every cell, against the dense arrays and the reference's layout
    int b[3];
    for (int d = 0; d < 3; d++) {
      b[d] = repo.blocksPerDimension(d);
      if (b[d] != n3[d] / 6 + 1 + (n3[d] % 6 ? 1 : 0)) return 6;
    }
    if (repo.numberOfVertices() != (unsigned)(b[0] * b[1] * b[2])) return 6;
    int64_t inner_total = 0;
    for (unsigned v = 8; v < 8 + repo.numberOfVertices(); v++) {
      BlockView bv = repo.loadVertex(v);
      const unsigned k = v - 8;
      const int bx = k % b[0], by = (k / b[0]) % b[1], bz = k / (b[0] * b[1]);
      for (int cell = 0; cell < LB_BLOCK_NUMBER_OF_CELLS; cell++) {
        const int x = cell % 6, y = (cell / 6) % 6, z = cell / 36;
        const int gx = 6 * bx + x - 3, gy = 6 * by + y - 3, gz = 6 * bz + z - 3;
        const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < n3[0] && gy < n3[1] && gz < n3[2];
        if ((bv.inner[cell] != 0) != in) return 7;
        if (!in) {
          for (int q = 0; q < LB_CURRENT_DIR; q++)
            if (bv.pdf[cell * LB_CURRENT_DIR + q] != 0.0) return 8;
          continue;
        }
        inner_total++;
        const int64_t i = gx + (int64_t)n3[0] * (gy + (int64_t)n3[1] * gz);
        for (int q = 0; q < LB_CURRENT_DIR; q++)
          if (bv.pdf[cell * LB_CURRENT_DIR + q] != f[(size_t)q * ncell + i]) return 9;
        if (bv.density[cell] != rho[i]) return 10;
      }
    }
    if (inner_total != ncell) return 11;
    bool threw = false;
    try {
      repo.loadVertex(3);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    if (!threw) return 12;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::printf("peano api ok\n");
  return 0;
}
