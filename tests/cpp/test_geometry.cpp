// Host-only driver of the block store's geometry flags (madpac_amd/csrc/peano_geometry.h), for the
// CPU tests: writes the inner and boundary flags of every block-store slot of a configured domain.
//
// usage: test_geometry out.bin nx ny nz dx ox oy oz [kind inverted p0 p1 p2 p3 p4 p5]...
//   kind 0 = hexahedron (p0..2 offset, p3..5 width), 1 = sphere (p0..2 centre, p3 radius)
// out.bin: int32 sx, sy, sz, then uint8 inner[sz][sy][sx], uint8 boundary[sz][sy][sx]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "peano_geometry.h"

using namespace madpac_amd::geometry;

int main(int argc, char** argv) {
  if (argc < 9 || (argc - 9) % 8 != 0) {
    std::fprintf(stderr, "usage: %s out.bin nx ny nz dx ox oy oz [kind inverted p0..p5]...\n", argv[0]);
    return 2;
  }
  const int n[3] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
  const double dx = std::strtod(argv[5], nullptr);
  const Vec3 off{std::strtod(argv[6], nullptr), std::strtod(argv[7], nullptr), std::strtod(argv[8], nullptr)};
  Domain g;
  for (int a = 9; a < argc; a += 8) {
    const int kind = std::atoi(argv[a]);
    const bool inv = std::atoi(argv[a + 1]) != 0;
    double p[6];
    for (int k = 0; k < 6; k++) p[k] = std::strtod(argv[a + 2 + k], nullptr);
    if (kind == 0) g.objects.push_back(Object::hexahedron({p[0], p[1], p[2]}, {p[3], p[4], p[5]}, inv));
    else g.objects.push_back(Object::sphere({p[0], p[1], p[2]}, p[3], inv));
  }
  BlockGrid grid(n, dx, off);
  std::vector<uint8_t> inner, boundary;
  grid.storeFlags(g, inner, boundary);
  FILE* fp = std::fopen(argv[1], "wb");
  if (!fp) return 3;
  const int s[3] = {6 * grid.blocks[0], 6 * grid.blocks[1], 6 * grid.blocks[2]};
  std::fwrite(s, sizeof(int), 3, fp);
  std::fwrite(inner.data(), 1, inner.size(), fp);
  std::fwrite(boundary.data(), 1, boundary.size(), fp);
  return std::fclose(fp) == 0 ? 0 : 4;
}
