// Cell-kernel fixture harness for the reference Peano block-LB (this container only).
//
// Calls the reference's per-block kernels directly on one 6^3 block of seeded,
// far-from-equilibrium populations and writes inputs and outputs:
//   BlockComputeCellValues::computeDensity / computeVelocity / computeFeq
//     (blocklatticeboltzmann/BlockComputeCellValues.cpp:9-68)
//   BlockBGK::collide via CollisionModelWrapper (collisionmodels/BlockBGK.cpp:18-45,
//     CollisionModelWrapper.cpp:28-39), which skips non-inner cells
//   BlockMovingWall::treatBoundaryAfterCollision (boundaryconditions/BlockMovingWall.cpp:13-33)
//   AABlockStreaming::streamingLocal + swapPdfsInsideCells
//     (nonlocaltreatment/AABlockStreaming.cpp:19-32,80-92)
//
// Output layout (little endian), all arrays AoS over the block (cell = x + 6y + 36z):
//   int32 magic 0x4C42434B, int32 seed
//   f64 tau_eff, f64 uw[3]
//   uint8 inner[216]
//   f64 pdf_in[216*19], f64 density[216], f64 velocity[648], f64 feq[216*19]
//   f64 pdf_collided[216*19]
//   int32 nlinks, then nlinks x (int32 cell, int32 q_incoming)  -- moving-wall links used
//   f64 pdf_after_wall[216*19]
//   f64 pdf_after_local_stream_and_swap[216*19]
#include "peano/applications/latticeboltzmann/LatticeBoltzmannDefinitions.h"
#include "peano/applications/latticeboltzmann/BoundaryData.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/BlockComputeCellValues.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/collisionmodels/BlockBGK.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/collisionmodels/CollisionModelWrapper.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/boundaryconditions/BlockMovingWall.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/nonlocaltreatment/AABlockStreaming.h"

#include <bitset>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace blb = peano::applications::latticeboltzmann::blocklatticeboltzmann;

// 64-bit LCG (Knuth MMIX constants), top 53 bits -> [0,1)
static unsigned long long s_state = 0;
static double urand() {
  s_state = s_state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(s_state >> 11) * (1.0 / 9007199254740992.0);
}

template <class T>
static void put(FILE* o, const T& v) { std::fwrite(&v, sizeof(T), 1, o); }

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s seed tau out.bin\n", argv[0]);
    return 2;
  }
  const int seed = std::atoi(argv[1]);
  const double tau = std::atof(argv[2]);
  s_state = (unsigned long long)seed * 0x9E3779B97F4A7C15ULL + 1ULL;
  const int NC = LB_BLOCK_NUMBER_OF_CELLS, Q = LB_CURRENT_DIR;

  std::bitset<LB_BLOCK_NUMBER_OF_CELLS> inner;
  tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, double> pdf;
  tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS, double> density(0.0);
  tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * DIMENSIONS, double> velocity(0.0);
  for (int i = 0; i < NC; i++) {
    inner[i] = urand() < 0.8;
    for (int q = 0; q < Q; q++) {
      // weights times a perturbation of up to +-40 %, plus a drift along c
      const double w = peano::applications::latticeboltzmann::LATTICEWEIGHTS._entries[q];
      pdf(i * Q + q) = inner[i] ? w * (0.6 + 0.8 * urand()) : 0.0;
    }
  }
  const tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, double> pdfIn(pdf);

  FILE* o = std::fopen(argv[3], "wb");
  if (!o) return 3;
  put(o, (int)0x4C42434B);
  put(o, seed);
  const double nu = 1.0 / 3.0 * (tau - 0.5);
  const double tauEff = 3.0 * nu + 0.5;
  put(o, tauEff);
  tarch::la::Vector<DIMENSIONS, double> uw(0.0);
  uw(0) = 0.05 + 0.1 * urand();
  uw(1) = -0.03 * urand();
  uw(2) = 0.0;
  for (int d = 0; d < DIMENSIONS; d++) put(o, uw(d));
  for (int i = 0; i < NC; i++) put(o, (unsigned char)(inner[i] ? 1 : 0));
  for (int k = 0; k < NC * Q; k++) put(o, pdfIn(k));

  blb::BLOCKCOMPUTECELLVALUES.computeDensity(inner, pdf, density);
  blb::BLOCKCOMPUTECELLVALUES.computeVelocity(inner, pdf, density, velocity);
  for (int i = 0; i < NC; i++) put(o, density(i));
  for (int k = 0; k < NC * DIMENSIONS; k++) put(o, velocity(k));
  for (int i = 0; i < NC; i++) {
    tarch::la::Vector<DIMENSIONS, double> u;
    for (int d = 0; d < DIMENSIONS; d++) u(d) = velocity(i * DIMENSIONS + d);
    tarch::la::Vector<LB_CURRENT_DIR, double> feq(0.0);
    blb::BLOCKCOMPUTECELLVALUES.computeFeq(density(i), u, feq);
    for (int q = 0; q < Q; q++) put(o, feq(q));
  }

  blb::collisionmodels::BlockBGK bgk(tauEff);
  blb::collisionmodels::CollisionModelWrapper wrapper(&bgk);
  wrapper.collide(inner, density, velocity, pdf);
  for (int k = 0; k < NC * Q; k++) put(o, pdf(k));

  // moving-wall links: every inner cell on the z = 5 face, every direction with c_z > 0
  blb::boundaryconditions::BlockMovingWall wall;
  peano::applications::latticeboltzmann::BoundaryData bd;
  bd.setVelocityL(uw);
  bd.setBoundaryType(peano::applications::latticeboltzmann::BoundaryData::MOVING_WALL_HW_BB);
  std::vector<tarch::la::Vector<2, unsigned int> > links;
  for (int i = 5 * 36; i < NC; i++) {
    if (!inner[i]) continue;
    for (int q = 0; q < Q; q++) {
      if (peano::applications::latticeboltzmann::LATTICEVELOCITIES._entries[q](2) > 0.0) {
        links.push_back(tarch::la::Vector<2, unsigned int>((unsigned int)i, (unsigned int)(Q - 1 - q)));
      }
    }
  }
  put(o, (int)links.size());
  for (size_t l = 0; l < links.size(); l++) {
    put(o, (int)links[l](0));
    put(o, (int)links[l](1));
    wall.treatBoundaryAfterCollision(0, bd, links[l], density, velocity, pdf);
  }
  for (int k = 0; k < NC * Q; k++) put(o, pdf(k));

  blb::nonlocaltreatment::AABlockStreaming streaming;
  streaming.streamingLocal(inner, pdf);
  streaming.swapPdfsInsideCells(pdf);
  for (int k = 0; k < NC * Q; k++) put(o, pdf(k));
  std::fclose(o);
  return 0;
}
