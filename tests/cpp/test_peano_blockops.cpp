// Drives the block solver's kernels of the reference-shaped C++ layer (madpac_amd/csrc/peano_blocklb.h:
// DeviceBlocks, BlockComputeCellValues, BlockBGK, BlockBoundaryConditionManager, AABlockStreaming) as
// a regular-grid solver mapping would (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:83-248), on
// the lid-driven cavity: per iteration the touchVertexFirstTime kernels on every vertex, the
// cross-block streaming of every Peano cell, the in-cell swap of every vertex, each kernel one call
// over all of them.
//
// usage: test_peano_blockops n tau ulid steps out.bin
// out.bin: f64 f[19][n][n][n] (SoA over the box, x fastest) after the steps
#include <array>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "peano_blocklb.h"

using namespace madpac_amd::blocklatticeboltzmann;

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s n tau ulid steps out.bin\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]), steps = std::atoi(argv[4]);
  const double tau = std::atof(argv[2]), ulid = std::atof(argv[3]);
  const int Q = LB_CURRENT_DIR, NC = LB_BLOCK_NUMBER_OF_CELLS, B = LB_BLOCKSIZE;
  // D3Q19, z-major (LB/LatticeBoltzmannDefinitions.cpp:113-151)
  static const int C[19][3] = {{0, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {0, 1, -1}, {-1, -1, 0}, {0, -1, 0},
                               {1, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {-1, 1, 0}, {0, 1, 0}, {1, 1, 0},
                               {0, -1, 1}, {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}};
  double w[19];
  for (int q = 0; q < Q; q++) {
    const int nz = (C[q][0] != 0) + (C[q][1] != 0) + (C[q][2] != 0);
    w[q] = nz == 0 ? 1.0 / 3.0 : (nz == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
  // tau_eff of the level (LB/configurations/LatticeBoltzmannConfiguration.cpp:247-269, LB/MultiLevelSimData.h:160)
  const double nu = 1.0 / 3.0 * (tau - 0.5);
  const double tauEff = 3.0 * nu + 0.5;
  // blocks per axis (BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:78-92); slot s of
  // block b holds cell 6 b + s - 3
  const int nb = n / B + 1 + (n % B ? 1 : 0);
  const int64_t nblocks = (int64_t)nb * nb * nb;
  DeviceBlocks store(nblocks);
  std::vector<int64_t> vertices;
  std::vector<lbm_block_boundary> links;
  auto cell_of = [&](int64_t v, int c, int d) {
    const int64_t b[3] = {v % nb, (v / nb) % nb, v / ((int64_t)nb * nb)};
    const int s[3] = {c % B, (c / B) % B, c / (B * B)};
    return (int)(B * b[d] + s[d] - 3);
  };
  for (int64_t v = 0; v < nblocks; v++) {
    BlockView b;
    b.inner.assign(NC, 0);
    b.pdf.resize(NC * Q);
    b.density.assign(NC, 1.0);
    b.velocity.assign(3 * NC, 0.0);
    for (int c = 0; c < NC; c++) {
      for (int q = 0; q < Q; q++) b.pdf[c * Q + q] = w[q];
      const int x = cell_of(v, c, 0), y = cell_of(v, c, 1), z = cell_of(v, c, 2);
      b.inner[c] = x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n;
      // the lid: top-layer links with c_z > 0 (SLB/BoundaryData4LidDrivenCavity.cpp:17-55)
      if (b.inner[c] && z == n - 1)
        for (int q = 0; q < Q; q++)
          if (C[q][2] > 0) {
            lbm_block_boundary l{};
            l.block = v;
            l.cell = c;
            l.q_in = Q - 1 - q;
            l.type = 1;
            l.u[0] = ulid;
            l.rho = 1.0;
            links.push_back(l);
          }
    }
    store.setBlock(v, b);
    vertices.push_back(v);
  }
  std::vector<std::array<int64_t, 8>> cells;
  for (int cz = 0; cz < nb - 1; cz++)
    for (int cy = 0; cy < nb - 1; cy++)
      for (int cx = 0; cx < nb - 1; cx++) {
        std::array<int64_t, 8> e{};
        for (int i = 0; i < 8; i++)
          e[i] = (cx + (i & 1)) + (int64_t)nb * ((cy + ((i >> 1) & 1)) + (int64_t)nb * (cz + (i >> 2)));
        cells.push_back(e);
      }

  const BlockComputeCellValues values;
  const BlockBGK bgk(tauEff);
  const BlockBoundaryConditionManager boundary;
  const AABlockStreaming streaming;
  for (int t = 0; t < steps; t++) {
    values.computeDensityAndVelocity(store, vertices);
    bgk.collide(store, vertices);
    boundary.treatBoundaryAfterCollision(store, links);
    streaming.streamingLocal(store, vertices);
    streaming.streamingNonLocal(store, cells);
    streaming.swapPdfsInsideCells(store, vertices);
  }

  const int64_t N = (int64_t)n * n * n;
  std::vector<double> f((size_t)(Q * N), 0.0);
  for (int64_t v = 0; v < nblocks; v++) {
    BlockView b;
    store.getBlock(v, b);
    for (int c = 0; c < NC; c++) {
      const int x = cell_of(v, c, 0), y = cell_of(v, c, 1), z = cell_of(v, c, 2);
      if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) continue;
      const int64_t i = x + (int64_t)n * (y + (int64_t)n * z);
      for (int q = 0; q < Q; q++) f[(size_t)(q * N + i)] = b.pdf[c * Q + q];
    }
  }
  FILE* o = std::fopen(argv[5], "wb");
  if (!o) return 1;
  std::fwrite(f.data(), sizeof(double), f.size(), o);
  std::fclose(o);
  return 0;
}
