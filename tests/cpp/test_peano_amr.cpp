// Drives BlockRestrictionProlongation of the reference-shaped C++ layer (madpac_amd/csrc/peano_blocklb.h)
// as the spacetree solver mapping calls it per fine vertex.
//
// usage: test_peano_amr in.bin out.bin
// in.bin: int32 ncases, then per case: int32 pos[3], int32 given (1: use the mask below instead of
//         findFineOverlapCells), uint8 mask[216], f64 coarse[4104], fine[4104], density[216], velocity[648]
// out.bin per case: int32 father[3], int32 subblock[3], uint8 overlap[216], f64 fine[4104],
//         density[216], velocity[648] after prolongateToFineLevel, f64 coarse[4104] after
//         restrictToCoarseLevel of the input fine block
#include <cstdio>
#include <vector>

#include "peano_blocklb.h"

using namespace madpac_amd::blocklatticeboltzmann;

template <class T>
static bool get(FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <class T>
static void put(FILE* f, const T* p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  FILE* out = std::fopen(argv[2], "wb");
  if (!in || !out) return 1;
  int ncases = 0;
  if (!get(in, &ncases, 1)) return 1;
  const int NC = LB_BLOCK_NUMBER_OF_CELLS, NP = NC * LB_CURRENT_DIR;
  BlockRestrictionProlongation rp;
  for (int k = 0; k < ncases; k++) {
    int pos[3], given = 0;
    std::vector<uint8_t> m(NC);
    std::vector<double> coarse(NP), fine(NP), density(NC), velocity(3 * NC);
    if (!get(in, pos, 3) || !get(in, &given, 1) || !get(in, m.data(), NC) || !get(in, coarse.data(), NP) ||
        !get(in, fine.data(), NP) || !get(in, density.data(), NC) || !get(in, velocity.data(), 3 * NC))
      return 1;
    const std::array<int, 3> p{pos[0], pos[1], pos[2]};
    const std::array<int, 3> father = rp.getFatherVertex(p), sub = rp.getSubBlock(p);
    std::vector<bool> overlap;
    if (given)
      overlap.assign(m.begin(), m.end());
    else
      rp.findFineOverlapCells(fine, overlap);
    std::vector<double> fine2(fine), coarse2(coarse);
    rp.prolongateToFineLevel(p, coarse, fine2, density, velocity, overlap);
    rp.restrictToCoarseLevel(p, coarse2, fine);
    put(out, father.data(), 3);
    put(out, sub.data(), 3);
    std::vector<uint8_t> ov(NC);
    for (int c = 0; c < NC; c++) ov[c] = overlap[c] ? 1 : 0;
    put(out, ov.data(), NC);
    put(out, fine2.data(), NP);
    put(out, density.data(), NC);
    put(out, velocity.data(), 3 * NC);
    put(out, coarse2.data(), NP);
  }
  std::fclose(in);
  return std::fclose(out) == 0 ? 0 : 1;
}
