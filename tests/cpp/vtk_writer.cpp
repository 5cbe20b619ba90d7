// Writes one block VTK file (madpac_amd::blocklatticeboltzmann::writeBlockVTK) from fields given
// on the host; no device is touched.  Usage:
//   vtk_writer nx ny nz dx fields.bin out.vtk
// fields.bin: uint8 fluid[n], f64 rho[n], f64 u[3n] (u interleaved), n = nx*ny*nz.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "peano_blocklb.h"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s nx ny nz dx fields.bin out.vtk\n", argv[0]);
    return 2;
  }
  const int n[3] = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
  const double dx = std::atof(argv[4]);
  const size_t cells = (size_t)n[0] * n[1] * n[2];
  std::vector<uint8_t> fluid(cells);
  std::vector<double> rho(cells), u(3 * cells);
  FILE* fp = std::fopen(argv[5], "rb");
  if (!fp) return 3;
  const bool ok = std::fread(fluid.data(), 1, cells, fp) == cells && std::fread(rho.data(), 8, cells, fp) == cells &&
                  std::fread(u.data(), 8, 3 * cells, fp) == 3 * cells;
  std::fclose(fp);
  if (!ok) return 4;
  const double offset[3] = {0.0, 0.0, 0.0};
  try {
    madpac_amd::blocklatticeboltzmann::writeBlockVTK(argv[6], n, dx, offset, fluid, rho, u);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 5;
  }
  return 0;
}
