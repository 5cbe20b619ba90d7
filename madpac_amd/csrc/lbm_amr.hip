// Restriction and prolongation between grid levels on a device-resident block store.
//
// Reference counterpart: BLB/BlockRestrictionProlongation.cpp, called per fine vertex from the
// spacetree solver mapping (BLB/mappings/SpacetreeGrid2RegularBlockSolver.cpp:160-240 in
// touchVertexFirstTime: overlap search, then prolongation after the collision; :340-366 in
// touchVertexLastTime: restriction).  A block is one Peano vertex's 6^3 cells in the reference's
// layout (populations AoS, pdf[19 * cell + q], cell = x + 6 y + 36 z; density[216]; velocity[648]),
// so blocks move between this store and GridManagementService-style host code unchanged.
//
// A refined coarse cell has 4^3 fine vertices at child positions pos in {0..3}^3.  The fine block at
// pos overlaps one sub-block (2^3 coarse cells) of its father vertex's block, each coarse cell
// covering 3^3 fine cells.  One call handles a batch of (coarse block, fine block, pos) links, one
// workgroup per link, so thousands of vertices of one level pair go in one launch:
//   * overlap search: the fine block's -1.0 markers (populations streamed in from temporary
//     vertices) name the neighbours that are temporary; each cell such a neighbour streams into is
//     an overlap cell.  The map from entries to neighbours is arithmetic: entry (cell, q) is filled
//     by the neighbour block holding position 6 + x - 3 c_q of the 3^3-block domain;
//   * prolongation: overlap cells take the coarse cell's populations, density and velocity;
//   * restriction: each coarse cell of the sub-block takes the mean of its 27 children per
//     population, unless a child carries the marker there.
// All three are HBM-bound byte movers (no arithmetic worth MFMA): the work per link is one fine
// block (32.8 KB of populations) plus one coarse sub-block.
#include <algorithm>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "lbm_blockstore.h"

namespace lbm {

// neighbour block (0..26, 13 = the block itself) whose populations fill entry (cell, q) after three
// streaming steps: the one holding position 6 + x - 3 c_q per axis
__device__ __forceinline__ int source_neighbour(int x, int y, int z, int q) {
  const int bx = (BS + x - 3 * CX(q)) / BS, by = (BS + y - 3 * CY(q)) / BS, bz = (BS + z - 3 * CZ(q)) / BS;
  return bx + 3 * by + 9 * bz;
}

// findFineOverlapCells (:32-79) for the fine block of each link; a new vertex is all markers and
// all overlap cells (SpacetreeGrid2RegularBlockSolver.cpp:186-193).  The block is staged through
// LDS in 16-byte pieces; reading it cell by cell from HBM (19 values 152 B apart per lane) fetched
// 4.5 x its bytes.
__global__ __launch_bounds__(256) void k_amr_overlap(const AmrLinkDev* __restrict__ links, int64_t nlinks,
                                                     double* __restrict__ pdf, uint8_t* __restrict__ mask) {
  const int64_t l = blockIdx.x;
  if (l >= nlinks) return;
  const AmrLinkDev L = links[l];
  double* f = pdf + L.fine * BP;
  uint8_t* m = mask + L.fine * BC;
  const int c = threadIdx.x;
  if (L.new_vertex) {
    double2* f2 = reinterpret_cast<double2*>(f);
    for (int k = c; k < BP2; k += 256) f2[k] = double2{-1.0, -1.0};
    if (c < BC) m[c] = 1;
    return;
  }
  __shared__ __align__(16) double s[BP];
  __shared__ unsigned temporary;
  stage_block(f, s, c);
  if (c == 0) temporary = 0u;
  __syncthreads();
  const int x = c % BS, y = (c / BS) % BS, z = c / (BS * BS);
  unsigned mine = 0u;  // neighbours this cell's entries come from
  if (c < BC) {
    unsigned marked = 0u;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int nb = source_neighbour(x, y, z, q);
      if (nb == 13) continue;
      mine |= 1u << nb;
      if (s[c * Q + q] == -1.0) marked |= 1u << nb;
    }
    if (marked) atomicOr(&temporary, marked);
  }
  __syncthreads();
  if (c < BC) m[c] = (mine & temporary) ? 1 : 0;
}

// prolongateToFineLevel (:83-160).  The 8 coarse cells of the sub-block and their density and
// velocity are computed once in LDS (each of the 27 fine cells above a coarse cell takes the same
// values: rho = sum v_q in q order, momentum per axis += v_q * c_q as a double product, u = m /
// rho); the populations go out as 16-byte pieces of the fine block that hold an overlap slot, a
// piece's non-overlap half rewritten with its own value.  (One thread per fine cell storing its 19
// values 152 B apart per lane wrote 5.6 x the algorithmic bytes.)
__global__ __launch_bounds__(256) void k_amr_prolongate(const AmrLinkDev* __restrict__ links, int64_t nlinks,
                                                        double* __restrict__ pdf, double* __restrict__ density,
                                                        double* __restrict__ velocity, const uint8_t* __restrict__ mask) {
  const int64_t l = blockIdx.x;
  if (l >= nlinks) return;
  const AmrLinkDev L = links[l];
  const int t = threadIdx.x;
  const int h = BS / 3;
  __shared__ double sc[8 * Q];
  __shared__ double srho[8], su[8 * 3];
  __shared__ uint8_t sm[BC];
  // both loads are issued before either is waited for (clamped lanes load an in-range value and
  // drop it)
  const int tc = min(t, 8 * Q - 1);
  const int k = tc / Q, q = tc % Q;  // coarse cell k of the sub-block: (k & 1, (k >> 1) & 1, k >> 2)
  const int cc = (L.sx * h + (k & 1)) + (L.sy * h + ((k >> 1) & 1)) * BS + (L.sz * h + (k >> 2)) * BS * BS;
  const double cv = pdf[L.coarse * BP + cc * Q + q];
  const uint8_t mv = mask[L.fine * BC + min(t, BC - 1)];
  if (t < 8 * Q) sc[t] = cv;
  if (t < BC) sm[t] = mv;
  __syncthreads();
  if (t < 8) {
    double rho = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const double v = sc[t * Q + q];
      rho += v;
      // the reference adds v * c_q per axis with c_q as a double (zero products included)
      m0 += v * (double)CX(q);
      m1 += v * (double)CY(q);
      m2 += v * (double)CZ(q);
    }
    srho[t] = rho;
    su[3 * t] = m0 / rho;
    su[3 * t + 1] = m1 / rho;
    su[3 * t + 2] = m2 / rho;
  }
  __syncthreads();
  // coarse cell of fine cell c: (x / 3, y / 3, z / 3) inside the sub-block
  auto coarse_of = [](int c) { return (c % BS) / 3 + 2 * (((c / BS) % BS) / 3) + 4 * ((c / (BS * BS)) / 3); };
  if (t < BC && sm[t]) {
    const int k = coarse_of(t);
    density[L.fine * BC + t] = srho[k];
    double* u = velocity + L.fine * BV + 3 * t;
    u[0] = su[3 * k];
    u[1] = su[3 * k + 1];
    u[2] = su[3 * k + 2];
  }
  double* f = pdf + L.fine * BP;
  double2* f2 = reinterpret_cast<double2*>(f);
  for (int p = t; p < BP2; p += 256) {
    const int s0 = 2 * p, s1 = 2 * p + 1, c0 = s0 / Q, c1 = s1 / Q;
    const bool o0 = sm[c0] != 0, o1 = sm[c1] != 0;
    if (!o0 && !o1) continue;
    const double v0 = o0 ? sc[coarse_of(c0) * Q + (s0 - c0 * Q)] : f[s0];
    const double v1 = o1 ? sc[coarse_of(c1) * Q + (s1 - c1 * Q)] : f[s1];
    f2[p] = double2{v0, v1};
  }
}

// restrictToCoarseLevel (:164-259): the fine block staged through LDS, then one thread per (coarse
// cell of the sub-block, population)
__global__ __launch_bounds__(256) void k_amr_restrict(const AmrLinkDev* __restrict__ links, int64_t nlinks,
                                                      double* __restrict__ pdf) {
  const int64_t l = blockIdx.x;
  if (l >= nlinks) return;
  const AmrLinkDev L = links[l];
  const int t = threadIdx.x;
  __shared__ __align__(16) double ff[BP];
  stage_block(pdf + L.fine * BP, ff, t);
  __syncthreads();
  if (t >= 8 * Q) return;
  const int q = t % Q, k = t / Q;
  const int x = k & 1, y = (k >> 1) & 1, z = k >> 2;
  const int h = BS / 3;
  const int cc = (L.sx * h + x) + (L.sy * h + y) * BS + (L.sz * h + z) * BS * BS;
  double r = 0.0;
  bool marked = false;
  for (int zz = 0; zz < 3; zz++)
    for (int yy = 0; yy < 3; yy++)
      for (int xx = 0; xx < 3; xx++) {
        const int fc = x * 3 + xx + (y * 3 + yy) * BS + (z * 3 + zz) * BS * BS;
        const double v = ff[fc * Q + q];
        marked = marked || v == -1.0;
        r += v;
      }
  double* cf = pdf + L.coarse * BP + cc * Q + q;
  const double mean = (1.0 / 27.0) * r;
  const double b = marked ? 1.0 : 0.0;
  // the reference's blend b * coarse + (1 - b) * mean, b = 1 or 0 (:246-248)
  *cf = b * *cf + (1.0 - b) * mean;
}

}  // namespace lbm

using namespace lbm;

namespace {
thread_local std::string g_blocks_create_error;

enum Op { OVERLAP, PROLONGATE, RESTRICT };

// Validates the links of one call and uploads them.  Within a call no block may be both a fine and
// a coarse block (the levels of one call are one coarse/fine pair), a link's blocks differ, fine
// blocks are unique (overlap search, prolongation) and coarse sub-blocks are unique (restriction),
// so no two workgroups write the same values.
int upload_links(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n, Op op) {
  if (n < 0 || (n > 0 && !links)) return bfail(bs, LBM_ERR_ARG, "link list is null");
  if (n > (int64_t)1 << 31) return bfail(bs, LBM_ERR_ARG, "too many links in one call");
  std::vector<AmrLinkDev> h((size_t)n);
  std::set<int64_t> fines, coarses;
  std::set<std::pair<int64_t, int>> subs;
  static const int sub_of[4] = {1, 2, 0, 1};  // createSubBlockVector (:280-288)
  for (int64_t k = 0; k < n; k++) {
    const lbm_amr_link& L = links[k];
    if (L.coarse < 0 || L.coarse >= bs->nblocks || L.fine < 0 || L.fine >= bs->nblocks)
      return bfail(bs, LBM_ERR_ARG, "link " + std::to_string(k) + ": block id out of range");
    if (L.coarse == L.fine) return bfail(bs, LBM_ERR_ARG, "link " + std::to_string(k) + ": coarse and fine block are the same");
    for (int d = 0; d < 3; d++)
      if (L.pos[d] < 0 || L.pos[d] > 3) return bfail(bs, LBM_ERR_ARG, "link " + std::to_string(k) + ": child position must be in 0..3");
    AmrLinkDev& D = h[(size_t)k];
    D.coarse = L.coarse;
    D.fine = L.fine;
    D.sx = sub_of[L.pos[0]];
    D.sy = sub_of[L.pos[1]];
    D.sz = sub_of[L.pos[2]];
    D.new_vertex = L.new_vertex ? 1 : 0;
    fines.insert(L.fine);
    coarses.insert(L.coarse);
    if (op == RESTRICT && !subs.insert({L.coarse, D.sx + 3 * D.sy + 9 * D.sz}).second)
      return bfail(bs, LBM_ERR_ARG, "link " + std::to_string(k) + ": two links restrict into the same coarse sub-block");
  }
  if (op != RESTRICT && (int64_t)fines.size() != n)
    return bfail(bs, LBM_ERR_ARG, "a fine block appears in two links");
  for (int64_t b : fines)
    if (coarses.count(b)) return bfail(bs, LBM_ERR_ARG, "a block is fine in one link and coarse in another (one level pair per call)");
  if (int rc = ensure_capacity(bs, &bs->d_links, &bs->cap_links, n, "link table")) return rc;
  if (n) BCHK(bs, hipMemcpyAsync(bs->d_links, h.data(), (size_t)n * sizeof(AmrLinkDev), hipMemcpyHostToDevice, bs->stream));
  return LBM_OK;
}

int run_links(lbm_blockstore* bs, int64_t n, Op op) {
  BCHK(bs, hipEventRecord(bs->ev0, bs->stream));
  if (n) {
    const dim3 grid((unsigned)n), block(256);
    if (op == OVERLAP)
      hipLaunchKernelGGL(k_amr_overlap, grid, block, 0, bs->stream, bs->d_links, n, bs->pdf, bs->mask);
    else if (op == PROLONGATE)
      hipLaunchKernelGGL(k_amr_prolongate, grid, block, 0, bs->stream, bs->d_links, n, bs->pdf, bs->density,
                         bs->velocity, bs->mask);
    else
      hipLaunchKernelGGL(k_amr_restrict, grid, block, 0, bs->stream, bs->d_links, n, bs->pdf);
    BCHK(bs, hipGetLastError());
  }
  BCHK(bs, hipEventRecord(bs->ev1, bs->stream));
  BCHK(bs, hipEventSynchronize(bs->ev1));
  BCHK(bs, hipEventElapsedTime(&bs->last_ms, bs->ev0, bs->ev1));
  return LBM_OK;
}
}  // namespace

extern "C" {

int lbm_blocks_create(lbm_blockstore** out, int64_t nblocks, int device) {
  if (!out) return LBM_ERR_ARG;
  *out = nullptr;
  lbm_blockstore* bs = new lbm_blockstore();
  auto bail = [&](int code, const std::string& m) {
    g_blocks_create_error = m;
    lbm_blocks_destroy(bs);
    return code;
  };
  if (nblocks < 1) return bail(LBM_ERR_ARG, "nblocks must be >= 1");
  bs->device = device;
  bs->nblocks = nblocks;
  if (hipSetDevice(device) != hipSuccess) return bail(LBM_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&bs->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&bs->ev0) != hipSuccess || hipEventCreate(&bs->ev1) != hipSuccess)
    return bail(LBM_ERR_HIP, "stream/event creation failed");
  const size_t nb = (size_t)nblocks;
  if (lbm::device_alloc(&bs->pdf, nb * BP * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&bs->density, nb * BC * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&bs->velocity, nb * BV * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&bs->mask, nb * BC) != hipSuccess || lbm::device_alloc(&bs->inner, nb * BC) != hipSuccess)
    return bail(LBM_ERR_NOMEM, "block store allocation failed");
  if (hipMemsetAsync(bs->pdf, 0, nb * BP * sizeof(double), bs->stream) != hipSuccess ||
      hipMemsetAsync(bs->density, 0, nb * BC * sizeof(double), bs->stream) != hipSuccess ||
      hipMemsetAsync(bs->velocity, 0, nb * BV * sizeof(double), bs->stream) != hipSuccess ||
      hipMemsetAsync(bs->mask, 0, nb * BC, bs->stream) != hipSuccess ||
      hipMemsetAsync(bs->inner, 1, nb * BC, bs->stream) != hipSuccess || hipStreamSynchronize(bs->stream) != hipSuccess)
    return bail(LBM_ERR_HIP, "block store clear failed");
  *out = bs;
  return LBM_OK;
}

int lbm_blocks_destroy(lbm_blockstore* bs) {
  if (!bs) return LBM_OK;
  if (bs->stream) (void)hipStreamSynchronize(bs->stream);
  void* ptrs[] = {bs->pdf, bs->density, bs->velocity, bs->mask, bs->inner, bs->d_links, bs->d_list, bs->d_cells,
                  bs->d_bnd, bs->d_off, bs->d_nbr, bs->d_valid, bs->d_xch, bs->d_pdfdiff, bs->d_bdens,
                  bs->boundary, bs->d_xfer, bs->d_vpos};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (bs->ev0) (void)hipEventDestroy(bs->ev0);
  if (bs->ev1) (void)hipEventDestroy(bs->ev1);
  if (bs->stream) (void)hipStreamDestroy(bs->stream);
  delete bs;
  return LBM_OK;
}

const char* lbm_blocks_last_error(const lbm_blockstore* bs) { return bs ? bs->err.c_str() : g_blocks_create_error.c_str(); }

int lbm_blocks_set(lbm_blockstore* bs, int64_t first, int64_t count, const double* pdf, const double* density,
                   const double* velocity) {
  if (!bs) return LBM_ERR_ARG;
  if (int rc = check_range(bs, first, count)) return rc;
  BCHK(bs, hipSetDevice(bs->device));
  if (pdf)
    BCHK(bs, hipMemcpyAsync(bs->pdf + first * BP, pdf, (size_t)count * BP * sizeof(double), hipMemcpyHostToDevice, bs->stream));
  if (density)
    BCHK(bs, hipMemcpyAsync(bs->density + first * BC, density, (size_t)count * BC * sizeof(double), hipMemcpyHostToDevice,
                            bs->stream));
  if (velocity)
    BCHK(bs, hipMemcpyAsync(bs->velocity + first * BV, velocity, (size_t)count * BV * sizeof(double), hipMemcpyHostToDevice,
                            bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

int lbm_blocks_get(lbm_blockstore* bs, int64_t first, int64_t count, double* pdf, double* density, double* velocity) {
  if (!bs) return LBM_ERR_ARG;
  if (int rc = check_range(bs, first, count)) return rc;
  BCHK(bs, hipSetDevice(bs->device));
  if (pdf)
    BCHK(bs, hipMemcpyAsync(pdf, bs->pdf + first * BP, (size_t)count * BP * sizeof(double), hipMemcpyDeviceToHost, bs->stream));
  if (density)
    BCHK(bs, hipMemcpyAsync(density, bs->density + first * BC, (size_t)count * BC * sizeof(double), hipMemcpyDeviceToHost,
                            bs->stream));
  if (velocity)
    BCHK(bs, hipMemcpyAsync(velocity, bs->velocity + first * BV, (size_t)count * BV * sizeof(double), hipMemcpyDeviceToHost,
                            bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

int lbm_blocks_set_overlap(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* overlap) {
  if (!bs) return LBM_ERR_ARG;
  if (!overlap) return bfail(bs, LBM_ERR_ARG, "overlap is null");
  if (int rc = check_range(bs, first, count)) return rc;
  BCHK(bs, hipSetDevice(bs->device));
  std::vector<uint8_t> m((size_t)count * BC);
  for (size_t k = 0; k < m.size(); k++) m[k] = overlap[k] ? 1 : 0;
  BCHK(bs, hipMemcpyAsync(bs->mask + first * BC, m.data(), m.size(), hipMemcpyHostToDevice, bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

int lbm_blocks_find_overlap(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n, uint8_t* overlap) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = upload_links(bs, links, n, OVERLAP)) return rc;
  if (int rc = run_links(bs, n, OVERLAP)) return rc;
  if (overlap && n) {
    std::vector<uint8_t> all((size_t)bs->nblocks * BC);
    BCHK(bs, hipMemcpyAsync(all.data(), bs->mask, all.size(), hipMemcpyDeviceToHost, bs->stream));
    BCHK(bs, hipStreamSynchronize(bs->stream));
    for (int64_t k = 0; k < n; k++) std::copy_n(all.data() + links[k].fine * BC, BC, overlap + k * BC);
  }
  return LBM_OK;
}

int lbm_blocks_prolongate(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = upload_links(bs, links, n, PROLONGATE)) return rc;
  return run_links(bs, n, PROLONGATE);
}

int lbm_blocks_restrict(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = upload_links(bs, links, n, RESTRICT)) return rc;
  return run_links(bs, n, RESTRICT);
}

float lbm_blocks_last_ms(const lbm_blockstore* bs) { return bs ? bs->last_ms : -1.f; }

}  // extern "C"
