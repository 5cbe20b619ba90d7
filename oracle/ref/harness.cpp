// Fixture-dump harness for the reference Peano block-LB (this container only).
//
// Runs the reference exactly as its own entry point does (madpac/main.cpp:57-128:
// fillLookupTables, parallel/shared-memory init, XML configuration, interpreteConfiguration)
// and afterwards writes every block held by the reference's block storage service
// (blocklatticeboltzmann/services/GridManagementService.h:74-86) to a raw binary file.
//
// The reference numbers its blocks 0..7 for the 2^D temporary vertices
// (GridManagementService.cpp:26-54) and from 8 upward in creation order
// (GridManagementService.cpp:58-153); the harness writes blocks 8 .. 8+nblocks-1.
//
// Output layout (little endian):
//   int32 magic 0x4C42524D, int32 nblocks, int32 cells_per_block (216), int32 q (19)
//   per block: uint8 inner[216], uint8 boundary[216], f64 density[216],
//              f64 velocity[216*3], f64 pdf[216*19]   (AoS, pdf[cell*19+q])
#include "peano/peano.h"
#include "tarch/configuration/ConfigurationRegistry.h"
#include "tarch/configuration/TopLevelConfiguration.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/services/GridManagementService.h"

#include <cstdio>
#include <cstdlib>
#include <list>
#include <string>

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s config.xml nblocks out.bin\n", argv[0]);
    return 2;
  }
  fillLookupTables();
  initMolecularDynamicsCoupling(&argc, &argv);
  if (initParallelEnvironment(&argc, &argv) != 0) return 3;
  if (initSharedMemoryEnvironment() != 0) return 4;

  std::list<tarch::configuration::TopLevelConfiguration*> configurations =
    tarch::configuration::ConfigurationRegistry::getInstance().readFile(argv[1], "peano-configuration");
  if (configurations.empty()) {
    std::fprintf(stderr, "configuration %s not found or invalid\n", argv[1]);
    return 5;
  }
  int rc = 0;
  for (std::list<tarch::configuration::TopLevelConfiguration*>::iterator p = configurations.begin();
       p != configurations.end() && rc == 0; ++p) {
    rc = (*p)->interpreteConfiguration();
  }
  if (rc != 0) return 6;

  const int nblocks = std::atoi(argv[2]);
  FILE* out = std::fopen(argv[3], "wb");
  if (!out) return 7;
  const int header[4] = {0x4C42524D, nblocks, LB_BLOCK_NUMBER_OF_CELLS, LB_CURRENT_DIR};
  std::fwrite(header, sizeof(int), 4, out);

  typedef peano::applications::latticeboltzmann::blocklatticeboltzmann::services::GridManagementService GMS;
  GMS& gms = GMS::getInstance();
  for (int b = 0; b < nblocks; b++) {
    gms.loadVertex(8 + b);
    if (gms.getVertexNumber() != (unsigned int)(8 + b)) {
      std::fprintf(stderr, "block %d missing from the block store\n", 8 + b);
      return 8;
    }
    unsigned char flags[LB_BLOCK_NUMBER_OF_CELLS];
    for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS; i++) flags[i] = gms.getInner()[i] ? 1 : 0;
    std::fwrite(flags, 1, LB_BLOCK_NUMBER_OF_CELLS, out);
    for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS; i++) flags[i] = gms.getBoundary()[i] ? 1 : 0;
    std::fwrite(flags, 1, LB_BLOCK_NUMBER_OF_CELLS, out);
    for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS; i++) {
      double v = gms.getDensity()(i);
      std::fwrite(&v, sizeof(double), 1, out);
    }
    for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS * DIMENSIONS; i++) {
      double v = gms.getVelocity()(i);
      std::fwrite(&v, sizeof(double), 1, out);
    }
    for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR; i++) {
      double v = gms.getPdf()(i);
      std::fwrite(&v, sizeof(double), 1, out);
    }
  }
  std::fclose(out);

  tarch::configuration::ConfigurationRegistry::getInstance().freeConfigurations(configurations);
  shutdownParallelEnvironment();
  shutdownSharedMemoryEnvironment();
  return 0;
}
