// Block-streaming fixture harness for the reference Peano block-LB (this container only).
//
// Dumps the reference's swap-streaming operators on 6^3 blocks (populations AoS, pdf[19 cell + q]):
//   CellTopology::getNonLocalStreamingIndex (nonlocaltreatment/CellTopology.cpp:435-465): the
//     (vertex1, vertex2, cell1, cell2, q) entries of one Peano cell's cross-block swaps;
//   AABlockStreaming::streamingLocal (AABlockStreaming.cpp:19-32) on a block of distinct values,
//     all cells inner and with seeded inner flags;
//   AABlockStreaming::streamingNonLocal (:37-53) on the 8 blocks of one Peano cell, likewise;
//   AABlockStreaming::swapPdfsInsideCells (:80-92).
// Values are pdf = 1 + k (k = 19 cell + q) + 10000 v for block v, so the outputs are permutations.
// Output (little endian): int32 magic 0x53545231, int32 seed, int32 nentries, nentries x int32[5],
//   then f64 local_all[4104], uint8 inner_local[216], f64 local_masked[4104], f64 swapped[4104],
//   f64 nonlocal_all[8][4104], uint8 inner_nonlocal[8][216], f64 nonlocal_masked[8][4104].
#include "peano/applications/latticeboltzmann/LatticeBoltzmannDefinitions.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/nonlocaltreatment/AABlockStreaming.h"
#include "peano/applications/latticeboltzmann/blocklatticeboltzmann/nonlocaltreatment/CellTopology.h"

#include <bitset>
#include <cstdio>
#include <cstdlib>

namespace nl = peano::applications::latticeboltzmann::blocklatticeboltzmann::nonlocaltreatment;

static unsigned long long s_state = 0;
static double urand() {
  s_state = s_state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(s_state >> 11) * (1.0 / 9007199254740992.0);
}
template <class T>
static void put(FILE* o, const T& v) { std::fwrite(&v, sizeof(T), 1, o); }

typedef tarch::la::Vector<LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR, double> Pdf;

static void fill(Pdf& p, int v) {
  for (int k = 0; k < LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR; k++) p(k) = 1.0 + k + 10000.0 * v;
}
static void dump(FILE* o, const Pdf& p) {
  for (int k = 0; k < LB_BLOCK_NUMBER_OF_CELLS * LB_CURRENT_DIR; k++) put(o, p(k));
}
static void mask(std::bitset<LB_BLOCK_NUMBER_OF_CELLS>& b, double pfluid) {
  for (int c = 0; c < LB_BLOCK_NUMBER_OF_CELLS; c++) b[c] = urand() < pfluid;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s seed out.bin\n", argv[0]);
    return 2;
  }
  const int seed = std::atoi(argv[1]);
  s_state = (unsigned long long)seed * 0x9E3779B97F4A7C15ULL + 1ULL;
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 1;
  put(o, (int)0x53545231);
  put(o, seed);
  nl::CellTopology topo;
  const std::vector<tarch::la::Vector<5, int> > idx = topo.getNonLocalStreamingIndex();
  put(o, (int)idx.size());
  for (unsigned k = 0; k < idx.size(); k++)
    for (int j = 0; j < 5; j++) put(o, (int)idx[k](j));

  nl::AABlockStreaming aa;
  std::bitset<LB_BLOCK_NUMBER_OF_CELLS> all;
  all.set();
  Pdf p;
  fill(p, 0);
  aa.streamingLocal(all, p);
  dump(o, p);
  std::bitset<LB_BLOCK_NUMBER_OF_CELLS> in;
  mask(in, 0.7);
  for (int c = 0; c < LB_BLOCK_NUMBER_OF_CELLS; c++) put(o, (unsigned char)(in[c] ? 1 : 0));
  fill(p, 0);
  aa.streamingLocal(in, p);
  dump(o, p);
  fill(p, 0);
  aa.swapPdfsInsideCells(p);
  dump(o, p);

  Pdf blocks[8];
  std::bitset<LB_BLOCK_NUMBER_OF_CELLS> inner[8];
  tarch::la::Vector<TWO_POWER_D, std::bitset<LB_BLOCK_NUMBER_OF_CELLS>*> ip;
  tarch::la::Vector<TWO_POWER_D, Pdf*> pp;
  for (int v = 0; v < 8; v++) {
    fill(blocks[v], v);
    inner[v].set();
    ip(v) = &inner[v];
    pp(v) = &blocks[v];
  }
  aa.streamingNonLocal(ip, pp);
  for (int v = 0; v < 8; v++) dump(o, blocks[v]);
  for (int v = 0; v < 8; v++) {
    fill(blocks[v], v);
    mask(inner[v], 0.7);
  }
  for (int v = 0; v < 8; v++)
    for (int c = 0; c < LB_BLOCK_NUMBER_OF_CELLS; c++) put(o, (unsigned char)(inner[v][c] ? 1 : 0));
  aa.streamingNonLocal(ip, pp);
  for (int v = 0; v < 8; v++) dump(o, blocks[v]);
  std::fclose(o);
  return 0;
}
