// Restriction/prolongation fixture harness for the reference Peano block-LB (this container only).
//
// Calls the reference's BlockRestrictionProlongation (blocklatticeboltzmann/
// BlockRestrictionProlongation.cpp) on seeded 6^3 blocks, for every child position
// (0..3)^3 of a fine vertex inside a refined coarse cell:
//   getFatherVertex, getSubBlock                              (:17-28, :264-276)
//   findFineOverlapCells on a fine block with -1.0 markers    (:32-62, overlap map :293-449)
//   prolongateToFineLevel with the overlap cells found         (:83-160)
//   restrictToCoarseLevel on the same fine block               (:164-259)
//
// Inputs come from a 64-bit LCG (Knuth MMIX constants, top 53 bits), drawn in this order per
// position: coarse pdf (0.5 + u), fine pdf (-1.0 with probability `pmark`, else 0.5 + u; for
// position index 21 every fine entry is -1.0, the state of a new persistent vertex), density
// (0.5 + u), velocity (u - 0.5).  Output (little endian), per position:
//   int32 father[3], int32 subblock[3], uint8 overlap[216],
//   f64 fine_after_prolongation[216*19], f64 density_after[216], f64 velocity_after[648],
//   f64 coarse_after_restriction[216*19]
// preceded by int32 magic 0x414D5231, int32 seed, f64 pmark.
#include "peano/applications/latticeboltzmann/LatticeBoltzmannDefinitions.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/BlockRestrictionProlongation.h"

#include <cstdio>
#include <cstdlib>

namespace blb = peano::applications::latticeboltzmann::blocklatticeboltzmann;

static unsigned long long s_state = 0;
static double urand() {
  s_state = s_state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(s_state >> 11) * (1.0 / 9007199254740992.0);
}

template <class T>
static void put(FILE* o, const T& v) { std::fwrite(&v, sizeof(T), 1, o); }

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s seed pmark out.bin\n", argv[0]);
    return 2;
  }
  const int seed = std::atoi(argv[1]);
  const double pmark = std::atof(argv[2]);
  s_state = (unsigned long long)seed * 0x9E3779B97F4A7C15ULL + 1ULL;
  const int NC = LB_BLOCK_NUMBER_OF_CELLS, Q = LB_CURRENT_DIR;
  FILE* o = std::fopen(argv[3], "wb");
  if (!o) return 1;
  put(o, (int)0x414D5231);
  put(o, seed);
  put(o, pmark);
  blb::BlockRestrictionProlongation rp;
  for (int p = 0; p < 64; p++) {
    tarch::la::Vector<DIMENSIONS, int> pos(p % 4, (p / 4) % 4, p / 16);
    tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, double> coarse, fine;
    tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS, double> density;
    tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * DIMENSIONS, double> velocity;
    for (int k = 0; k < NC * Q; k++) coarse(k) = 0.5 + urand();
    for (int k = 0; k < NC * Q; k++) {
      const double r = urand();
      const double v = 0.5 + urand();
      fine(k) = (p == 21 || r < pmark) ? -1.0 : v;
    }
    for (int k = 0; k < NC; k++) density(k) = 0.5 + urand();
    for (int k = 0; k < NC * DIMENSIONS; k++) velocity(k) = urand() - 0.5;

    const tarch::la::Vector<DIMENSIONS, int> father = rp.getFatherVertex(pos);
    const tarch::la::Vector<DIMENSIONS, int> sub = rp.getSubBlock(pos);
    tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS, bool> overlap(false);
    rp.findFineOverlapCells(fine, overlap);

    tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, double> fine2(fine), coarse2(coarse);
    rp.prolongateToFineLevel(pos, coarse, fine2, density, velocity, overlap);
    rp.restrictToCoarseLevel(pos, coarse2, fine);

    for (int d = 0; d < 3; d++) put(o, (int)father(d));
    for (int d = 0; d < 3; d++) put(o, (int)sub(d));
    for (int k = 0; k < NC; k++) put(o, (unsigned char)(overlap(k) ? 1 : 0));
    for (int k = 0; k < NC * Q; k++) put(o, fine2(k));
    for (int k = 0; k < NC; k++) put(o, density(k));
    for (int k = 0; k < NC * DIMENSIONS; k++) put(o, velocity(k));
    for (int k = 0; k < NC * Q; k++) put(o, coarse2(k));
  }
  std::fclose(o);
  return 0;
}
