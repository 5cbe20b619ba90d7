// The device-resident store of 6^3 vertex blocks shared by the grid-level transfer operators
// (lbm_amr.hip) and the block solver's per-block and per-Peano-cell operators (lbm_blockops.hip).
//
// A block is one Peano vertex's 6^3 cells in the layout of GridManagementService::VertexData
// (BLB/services/GridManagementService.h:74-86): populations AoS pdf[19 * cell + q], density[216],
// velocity[648] (u[3 * cell + d]), cell = x + 6 y + 36 z; the inner bitset is one byte per cell.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lbm_capi.h"
#include "lbm_cell.h"
#include "lbm_internal.h"

namespace lbm {

constexpr int BS = 6;                  // block edge
constexpr int BC = BS * BS * BS;       // cells per block
constexpr int BP = BC * Q;             // populations per block
constexpr int BV = BC * 3;             // velocity entries per block

struct AmrLinkDev {
  int64_t coarse, fine;
  int sx, sy, sz;    // sub-block of the coarse block (0..2 per axis)
  int new_vertex;
};

// A block's 4,104 populations in 2,052 16-byte pieces (a block starts on a 16-byte boundary): 8 rounds
// of 256 threads and 4 more.  All loads of a thread are issued before its first LDS write, so a
// block's 32.8 KB are in flight at once; stores go out the same way.  For 256-thread workgroups.
constexpr int BP2 = BP / 2, NFULL = BP2 / 256, TAIL = BP2 - 256 * NFULL;
static_assert(NFULL == 8 && TAIL == 4, "the staging below is written for 8 full rounds and 4 pieces");

// The staging in two halves, for kernels that issue other loads while a block's are in flight.
struct BlockPieces {
  double2 r0, r1, r2, r3, r4, r5, r6, r7, r8;
};
__device__ __forceinline__ BlockPieces load_block(const double* __restrict__ src, int t) {
  const double2* g = reinterpret_cast<const double2*>(src);
  BlockPieces p;
  p.r0 = g[t]; p.r1 = g[t + 256]; p.r2 = g[t + 512]; p.r3 = g[t + 768];
  p.r4 = g[t + 1024]; p.r5 = g[t + 1280]; p.r6 = g[t + 1536]; p.r7 = g[t + 1792];
  p.r8 = t < TAIL ? g[t + 2048] : double2{0.0, 0.0};
  return p;
}
__device__ __forceinline__ void put_block(const BlockPieces& p, double* s, int t) {
  double2* s2 = reinterpret_cast<double2*>(s);
  s2[t] = p.r0; s2[t + 256] = p.r1; s2[t + 512] = p.r2; s2[t + 768] = p.r3;
  s2[t + 1024] = p.r4; s2[t + 1280] = p.r5; s2[t + 1536] = p.r6; s2[t + 1792] = p.r7;
  if (t < TAIL) s2[t + 2048] = p.r8;
}
__device__ __forceinline__ void stage_block(const double* __restrict__ src, double* s, int t) {
  put_block(load_block(src, t), s, t);
}

__device__ __forceinline__ void store_block(const double* s, double* __restrict__ dst, int t) {
  const double2* s2 = reinterpret_cast<const double2*>(s);
  double2* g = reinterpret_cast<double2*>(dst);
  const double2 w0 = s2[t], w1 = s2[t + 256], w2 = s2[t + 512], w3 = s2[t + 768];
  const double2 w4 = s2[t + 1024], w5 = s2[t + 1280], w6 = s2[t + 1536], w7 = s2[t + 1792];
  g[t] = w0; g[t + 256] = w1; g[t + 512] = w2; g[t + 768] = w3;
  g[t + 1024] = w4; g[t + 1280] = w5; g[t + 1536] = w6; g[t + 1792] = w7;
  if (t < TAIL) g[t + 2048] = s2[t + 2048];
}

// one boundary link of a block, as lbm_block_boundary with the block's position in the call's list
struct BlockLinkDev {
  int64_t block;
  int32_t cell, q_in, type, pad;
  double u[3];
  double rho;
};
static_assert(sizeof(BlockLinkDev) == sizeof(lbm_block_boundary), "link layout");

}  // namespace lbm

struct lbm_blockstore {
  int device = 0;
  int64_t nblocks = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  double* pdf = nullptr;       // nblocks * 4104
  double* density = nullptr;   // nblocks * 216
  double* velocity = nullptr;  // nblocks * 648
  uint8_t* mask = nullptr;     // nblocks * 216 overlap cells of the last search (hasMinusOneEntries)
  uint8_t* inner = nullptr;    // nblocks * 216 inner flags (1 = inner)
  lbm::AmrLinkDev* d_links = nullptr;
  int64_t cap_links = 0;
  int64_t* d_list = nullptr;   // block list, or 8 blocks per Peano cell
  int64_t cap_list = 0;
  int64_t* d_cells = nullptr;  // Peano cells of lbm_blocks_iterate
  int64_t cap_cells = 0;
  lbm::BlockLinkDev* d_bnd = nullptr;  // boundary links grouped by list position
  int64_t cap_bnd = 0;
  int32_t* d_off = nullptr;    // per list position: first link (n + 1 entries)
  int64_t cap_off = 0;
  int64_t* d_nbr = nullptr;    // lbm_blocks_iterate: 18 neighbours per block
  int64_t cap_nbr = 0;
  uint32_t* d_valid = nullptr; // lbm_blocks_iterate: 32 swap-mask words per block
  int64_t cap_valid = 0;
  uint8_t* d_vpos = nullptr;   // lbm_blocks_iterate: per block, the vertex positions it has in listed cells
  int64_t cap_vpos = 0;
  double* d_xch = nullptr;     // lbm_blocks_iterate: two exchange arrays of 1,008 values per block
  int64_t cap_xch = 0;
  double* d_pdfdiff = nullptr; // exchange between ranks: 1,008 differences per block (allocated on first use)
  double* d_bdens = nullptr;   // exchange between ranks: 208 boundary densities per block
  uint8_t* boundary = nullptr; // exchange between ranks: nblocks * 216 geometry-boundary flags
  double* d_xfer = nullptr;    // exchange between ranks: staging of the values a call copies in or out
  int64_t cap_xfer = 0;
  uint64_t inner_gen = 0;      // bumped by every lbm_blocks_set_inner
  std::vector<int64_t> iter_cells;  // the cell list d_nbr and d_valid were built for ...
  uint64_t iter_inner_gen = ~0ull;  // ... and the inner flags
  float last_ms = 0.f;
};

namespace lbm {

inline int bfail(lbm_blockstore* bs, int code, const std::string& m) {
  if (bs) bs->err = m;
  return code;
}

#define BCHK(bs, call)                                                                                         \
  do {                                                                                                         \
    hipError_t e_ = (call);                                                                                    \
    if (e_ != hipSuccess) return lbm::bfail(bs, LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int check_range(lbm_blockstore* bs, int64_t first, int64_t count) {
  if (first < 0 || count < 0 || first > bs->nblocks || count > bs->nblocks - first)
    return bfail(bs, LBM_ERR_ARG, "block range out of bounds");
  return LBM_OK;
}

// grows a device array to at least n elements (contents are not kept)
template <class T>
int ensure_capacity(lbm_blockstore* bs, T** p, int64_t* cap, int64_t n, const char* what) {
  if (n <= *cap) return LBM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (device_alloc(p, (size_t)n * sizeof(T)) != hipSuccess)
    return bfail(bs, LBM_ERR_NOMEM, std::string(what) + " allocation failed");
  *cap = n;
  return LBM_OK;
}

}  // namespace lbm
