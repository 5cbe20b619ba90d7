// Block-boundary index harness for the reference Peano block-LB (this container only).
//
// Dumps the two index tables the reference's vertex-overlap exchange between ranks is built on:
//   parallel::BLOCKBOUNDARYINDEX4PDFS (parallel/BlockBoundaryIndex4Pdfs.cpp:13-77): for each of the
//     LB_PDFS_ON_BLOCKBOUNDARY populations that leave a 6^3 block, the slot of the population that
//     enters the block there (getIncomingPdf) and of the one that leaves it (getOutgoingPdf);
//   BLOCKBOUNDARYINDEX (BlockBoundaryIndex.cpp:15-88): the block index of each of the
//     LB_BLOCK_NUMBER_OF_CELLS_ON_BLOCKBOUNDARY cells near the block boundary (getBlockIndex).
// Output (little endian): int32 magic 0x50415231, int32 npdf, npdf x int32[2] (incoming, outgoing),
//   int32 ncells, ncells x int32.
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/BlockBoundaryIndex.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/parallel/BlockBoundaryIndex4Pdfs.h"

#include <cstdio>

namespace blb = peano::applications::latticeboltzmann::blocklatticeboltzmann;

template <class T>
static void put(FILE* o, const T& v) { std::fwrite(&v, sizeof(T), 1, o); }

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  FILE* o = std::fopen(argv[1], "wb");
  if (!o) return 1;
  put(o, (int)0x50415231);
  put(o, (int)LB_PDFS_ON_BLOCKBOUNDARY);
  for (int i = 0; i < LB_PDFS_ON_BLOCKBOUNDARY; i++) {
    put(o, (int)blb::parallel::BLOCKBOUNDARYINDEX4PDFS.getIncomingPdf(i));
    put(o, (int)blb::parallel::BLOCKBOUNDARYINDEX4PDFS.getOutgoingPdf(i));
  }
  put(o, (int)LB_BLOCK_NUMBER_OF_CELLS_ON_BLOCKBOUNDARY);
  for (int i = 0; i < LB_BLOCK_NUMBER_OF_CELLS_ON_BLOCKBOUNDARY; i++)
    put(o, (int)blb::BLOCKBOUNDARYINDEX.getBlockIndex(i));
  std::fclose(o);
  return 0;
}
