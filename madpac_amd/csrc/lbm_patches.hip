// Batched uneven patches: thousands of cubic patches (edge n, one-cell ghost layer) advanced by
// one persistent launch per step, with the ghost-layer exchange fused into the producer.
//
// Reference counterpart: each Peano vertex owns one regular block, and one iteration runs the
// per-block collision (touchVertexFirstTime), the streaming between the blocks that meet in a
// Peano cell (handleCell -> AABlockStreaming::streamingNonLocal, the reference's "ghost-layer
// exchange") and the in-cell swap (touchVertexLastTime); BLB/mappings/RegularGrid2RegularBlockSolver.cpp
// :83-248, BLB/nonlocaltreatment/AABlockStreaming.cpp:37-53.  Here a patch is stored with a
// ghost layer: the state is post-collision, f_q(x) = s_q(x - c_q) for every interior cell, and a
// cell on a patch boundary, after colliding, also writes each population that leaves the patch
// into the ghost layer of the neighbouring patch (face or edge neighbour, a 19-entry table per
// patch), where the next step's pull finds it.  No grid-wide barrier is needed inside a step: the
// kernel boundary orders the steps.
//
// Work decomposition: patches are cut into tiles of tile_x * tile_y * tile_z cells; a persistent
// grid of (CUs * blocks_per_cu) workgroups walks the tile list (largest patches first) with a
// fixed stride, so every workgroup exits after its last tile.  lbm_patches_autotune times
// candidate tilings on the device and keeps the fastest.  Steps can be captured once in a
// hipGraph (two launches: A -> B, B -> A) and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_capi.h"
#include "lbm_cell.h"
#include "lbm_internal.h"

namespace lbm {

struct PatchDev {
  int64_t base;      // element offset of the patch: 19 interior planes of n^3, then its ghost records
  int n;             // interior edge
  int pad_;
  double omega;      // 1 / tau_eff of the patch
  int nbr[Q];        // neighbour patch in direction c_q
  int64_t nbase[Q];  // that neighbour's base (one load fewer on the ghost path)
  int64_t cbase;     // offset of the patch's first interior cell in per-cell arrays (flag words, user arrays)
};

// Everything one tile iteration needs, so that a workgroup fetches a single record per tile (and
// the next one while it works on the current).
struct TileDev {
  int64_t base;      // the patch's base
  double omega;
  int patch;
  int n;
  int x0, y0, z0;    // tile origin (interior coordinates), or the first cell's linear index
};

// Storage of one patch (element offsets from its base):
//   19 interior planes of n^3 cells, x fastest: plane q holds the post-collision population q;
//   then the ghost records of the populations that enter the patch from outside:
//   6 faces (axis a, side e = -1/+1; id 2a + (e > 0)) x the 5 populations with c_a = -e, each an
//   n x n array indexed by the two other coordinates of the source cell (lower axis fastest);
//   12 edges (axes a < b, sides e_a, e_b; id 4 pair + 2 (e_a > 0) + (e_b > 0), pair xy 0, xz 1,
//   yz 2) x the one population c = -(e_a, e_b), an n-array indexed by the remaining coordinate.
// The neighbour across the face or edge writes a ghost record when it streams out; the patch reads
// it when it pulls.  Interior rows and planes are contiguous, and the ghost records of a face are
// contiguous along its lower in-face axis, so both sides access them with consecutive lanes.
__host__ __device__ inline int64_t patch_elems(int n) {
  const int64_t n1 = n;
  return ((int64_t)Q * n1 * n1 * n1 + 30 * n1 * n1 + 12 * n1 + 63) / 64 * 64;
}

// index of q among the 5 populations that cross a face normal to axis a (c_{q,a} != 0)
__host__ __device__ constexpr int face_local(int q, int a) {
  const int cb = a == 0 ? CY(q) : CX(q);  // the two other components, lower axis first
  const int cc = a == 2 ? CY(q) : CZ(q);
  return (cb == 0 && cc == 0) ? 0 : (cb != 0 ? (cb < 0 ? 1 : 2) : (cc < 0 ? 3 : 4));
}

// Offset of slot s = (s0, s1, s2) of population q: the own plane when s is interior, else the
// ghost record filled by the neighbour that owns s (s outside along one axis: a face; two: an edge).
__host__ __device__ __forceinline__ int64_t slot_offset(int q, int n, int s0, int s1, int s2) {
  const int64_t n1 = n, n2 = n1 * n1, n3 = n2 * n1;
  const int e0 = s0 < 0 ? -1 : (s0 >= n ? 1 : 0);
  const int e1 = s1 < 0 ? -1 : (s1 >= n ? 1 : 0);
  const int e2 = s2 < 0 ? -1 : (s2 >= n ? 1 : 0);
  if ((e0 | e1 | e2) == 0) return (int64_t)q * n3 + s0 + n1 * (s1 + n1 * s2);
  const int64_t G0 = (int64_t)Q * n3;
  if (e1 == 0 && e2 == 0) return G0 + (5 * (0 + (e0 > 0)) + face_local(q, 0)) * n2 + s1 + n1 * s2;
  if (e0 == 0 && e2 == 0) return G0 + (5 * (2 + (e1 > 0)) + face_local(q, 1)) * n2 + s0 + n1 * s2;
  if (e0 == 0 && e1 == 0) return G0 + (5 * (4 + (e2 > 0)) + face_local(q, 2)) * n2 + s0 + n1 * s1;
  const int64_t E0 = G0 + 30 * n2;
  if (e2 == 0) return E0 + (0 + 2 * (e0 > 0) + (e1 > 0)) * n1 + s2;
  if (e1 == 0) return E0 + (4 + 2 * (e0 > 0) + (e2 > 0)) * n1 + s1;
  return E0 + (8 + 2 * (e1 > 0) + (e2 > 0)) * n1 + s0;
}

// D3Q19 index of a face or edge direction e (at most two non-zero components), by arithmetic on
// the z-major ordering: z = 0 holds the 3x3 block 5..13, z = -1 / +1 the five directions 0..4 / 14..18.
__host__ __device__ constexpr int dir_index(int ex, int ey, int ez) {
  return ez == 0 ? 9 + ex + 3 * ey : (ez < 0 ? 2 : 16) + ex + 2 * ey;
}
constexpr bool dir_index_matches_table() {
  for (int q = 0; q < Q; q++)
    if (dir_index(CX(q), CY(q), CZ(q)) != q) return false;
  return true;
}
static_assert(dir_index_matches_table(), "dir_index must invert the D3Q19 ordering");

// M32: FP32 arithmetic for FP32 storage (else the hybrid arithmetic of collide_f32s_h)
template <typename T, bool M32>
__device__ __forceinline__ void collide_patch(T (&f)[Q], double omega) {
  if constexpr (sizeof(T) == 8) {
    CollideParams p{1.0 / omega, omega, {0.0, 0.0, 0.0}};
    double u[3];
    collide_f64<false, false>(f, p, 0.0, u);
  } else if constexpr (M32) {
    CollideParams32 p{(float)omega, 0.f, {0.f, 0.f, 0.f}};
    float u[4];
    collide_f32s<false, false>(f, p, 0.f, u);
  } else {
    CollideParams p{1.0 / omega, omega, {0.0, 0.0, 0.0}};
    CollideParams32 p32{(float)omega, 0.f, {0.f, 0.f, 0.f}};
    float u[4];
    collide_f32s_h<false, false>(f, p, p32, 0.f, u);
  }
}

// Slot arithmetic of one interior cell without branches (the step kernels' form of slot_offset).
// Along each axis a cell touches at most one face (n >= 2).  The pull source x - c_q lies outside
// along axis a iff c_{q,a} = +1 at coordinate 0 or -1 at n - 1; the push target x + c_q iff
// c_{q,a} = +1 at n - 1 or -1 at 0.  Every alternative (own plane, face record, edge record) is a
// tile-uniform base plus a per-lane term, picked by those flags; all offsets fit 32 bits (n <= 256).
struct PatchLane {
  int x, y, z, i, n, n2, n3;
  bool lx, hx, ly, hy, lz, hz;  // coordinate == 0 / == n - 1
};

__device__ __forceinline__ PatchLane patch_lane(int n, int x, int y, int z) {
  PatchLane L;
  L.x = x; L.y = y; L.z = z;
  L.n = n; L.n2 = n * n; L.n3 = n * n * n;
  L.i = x + n * (y + n * z);
  L.lx = x == 0; L.hx = x == n - 1;
  L.ly = y == 0; L.hy = y == n - 1;
  L.lz = z == 0; L.hz = z == n - 1;
  return L;
}

__device__ __forceinline__ bool src_outside(int c, bool lo, bool hi) { return c > 0 ? lo : (c < 0 ? hi : false); }
__device__ __forceinline__ bool dst_outside(int c, bool lo, bool hi) { return c > 0 ? hi : (c < 0 ? lo : false); }

// Offset from the patch base of the slot that population q of the cell is pulled from: equal to
// slot_offset(q, n, x - c_x, y - c_y, z - c_z).
__device__ __forceinline__ int pull_slot(int q, const PatchLane& L) {
  const int cx = CX(q), cy = CY(q), cz = CZ(q);
  const bool ox = src_outside(cx, L.lx, L.hx), oy = src_outside(cy, L.ly, L.hy), oz = src_outside(cz, L.lz, L.hz);
  const int n = L.n, n2 = L.n2, G0 = Q * L.n3, E0 = G0 + 30 * n2;
  int off = q * L.n3 + L.i - (cx + n * cy + n2 * cz);
  // a face record holds the source's two other coordinates (lower axis fastest), an edge record the third
  if (cx != 0 && ox) off = G0 + (5 * (cx < 0) + face_local(q, 0)) * n2 + (L.y - cy) + n * (L.z - cz);
  if (cy != 0 && oy) off = G0 + (5 * (2 + (cy < 0)) + face_local(q, 1)) * n2 + (L.x - cx) + n * (L.z - cz);
  if (cz != 0 && oz) off = G0 + (5 * (4 + (cz < 0)) + face_local(q, 2)) * n2 + (L.x - cx) + n * (L.y - cy);
  if (cx != 0 && cy != 0 && ox && oy) off = E0 + (2 * (cx < 0) + (cy < 0)) * n + L.z;
  if (cx != 0 && cz != 0 && ox && oz) off = E0 + (4 + 2 * (cx < 0) + (cz < 0)) * n + L.y;
  if (cy != 0 && cz != 0 && oy && oz) off = E0 + (8 + 2 * (cy < 0) + (cz < 0)) * n + L.x;
  return off;
}

// One interior cell (x, y, z) of a patch: pull (ghost records included), collide, store the own
// planes, and write each population that leaves the patch into the receiving neighbour's ghost
// record.  The receiver (direction e = the crossed face or edge, e = c_q on the crossed axes) sees
// the source at s = (x, y, z) - n e, so the record is indexed by the cell's own other coordinates:
// the write goes to slot_offset(q, n, x - n e_x, y - n e_y, z - n e_z) of that neighbour.  The
// neighbours' bases are tile-uniform loads from the patch record.
// FL: the patch set carries fluid flags (the reference block's inner bitset plus its boundary links,
// BLB/services/GridManagementService.h:74-80, SLB/services/ReceiveBoundaryDataService.cpp:107-185),
// one word per interior cell as in the dense lattice: bit 31 fluid, bit q set when x + c_q is not
// fluid.  A solid cell does nothing.  Across a bounce-back link (x - c_q not fluid) population q is
// the cell's own outgoing population 18 - q of the previous step (halfway bounce-back,
// BLB/boundaryconditions/BlockHalfwayBounceBack.cpp:12-27), and no ghost record is written toward a
// solid cell (no fluid cell pulls from it).
template <typename T, bool M32, bool FL>
__device__ __forceinline__ void patch_cell(const PatchDev* __restrict__ pp, const uint32_t* __restrict__ words,
                                           const T* __restrict__ s, T* __restrict__ d, T* __restrict__ dst, int n,
                                           double omega, int x, int y, int z) {
  T f[Q];
  uint32_t w = 0u;
  {
    const PatchLane L = patch_lane(n, x, y, z);
    if (FL) {
      w = words[pp->cbase + L.i];
      if (!(w & FLUID)) return;
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const bool bounce = FL && ((w >> (Q - 1 - q)) & 1u);
      f[q] = s[bounce ? (Q - 1 - q) * L.n3 + L.i : pull_slot(q, L)];
    }
  }
  collide_patch<T, M32>(f, omega);
  // Rebuild the lane terms from opaque copies of the coordinates: kept live through the collision
  // they cost more registers than the collision leaves (spills).
  asm volatile("" : "+v"(x), "+v"(y), "+v"(z));
  const PatchLane L = patch_lane(n, x, y, z);
#pragma unroll
  for (int q = 0; q < Q; q++) __builtin_nontemporal_store(f[q], &d[q * L.n3 + L.i]);
  if (L.lx | L.hx | L.ly | L.hy | L.lz | L.hz) {
    const int n2 = L.n2, G0 = Q * L.n3, E0 = G0 + 30 * n2;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      if (q == REST) continue;
      const int cx = CX(q), cy = CY(q), cz = CZ(q);
      const bool ox = dst_outside(cx, L.lx, L.hx), oy = dst_outside(cy, L.ly, L.hy), oz = dst_outside(cz, L.lz, L.hz);
      if (!(ox | oy | oz)) continue;
      if (FL && ((w >> q) & 1u)) continue;
      int64_t nb = 0;
      int off = 0;
      if (cx != 0 && ox) { nb = pp->nbase[dir_index(cx, 0, 0)]; off = G0 + (5 * (cx < 0) + face_local(q, 0)) * n2 + L.y + n * L.z; }
      if (cy != 0 && oy) { nb = pp->nbase[dir_index(0, cy, 0)]; off = G0 + (5 * (2 + (cy < 0)) + face_local(q, 1)) * n2 + L.x + n * L.z; }
      if (cz != 0 && oz) { nb = pp->nbase[dir_index(0, 0, cz)]; off = G0 + (5 * (4 + (cz < 0)) + face_local(q, 2)) * n2 + L.x + n * L.y; }
      if (cx != 0 && cy != 0 && ox && oy) { nb = pp->nbase[q]; off = E0 + (2 * (cx < 0) + (cy < 0)) * n + L.z; }
      if (cx != 0 && cz != 0 && ox && oz) { nb = pp->nbase[q]; off = E0 + (4 + 2 * (cx < 0) + (cz < 0)) * n + L.y; }
      if (cy != 0 && cz != 0 && oy && oz) { nb = pp->nbase[q]; off = E0 + (8 + 2 * (cy < 0) + (cz < 0)) * n + L.x; }
      dst[nb + off] = f[q];
    }
  }
}

// Box tiles of TX x TY x TZ cells (cells outside the patch idle).
// The tiles of workgroup blockIdx.x: strided (t = w, w + G, ...) or, with CHUNK, the contiguous
// range [w T / G, (w + 1) T / G).
template <bool CHUNK>
__device__ __forceinline__ void tile_range(int ntiles, int& first, int& end, int& stride) {
  if (CHUNK) {
    first = (int)((int64_t)blockIdx.x * ntiles / gridDim.x);
    end = (int)(((int64_t)blockIdx.x + 1) * ntiles / gridDim.x);
    stride = 1;
  } else {
    first = blockIdx.x;
    end = ntiles;
    stride = gridDim.x;
  }
}

template <typename T, int TX, int TY, int TZ, bool M32, bool FL, bool CHUNK>
__global__ __launch_bounds__(256) void k_patch_step(const PatchDev* __restrict__ patches,
                                                    const TileDev* __restrict__ tiles, int ntiles,
                                                    const uint32_t* __restrict__ words,
                                                    const T* __restrict__ src, T* __restrict__ dst) {
  constexpr int TC = TX * TY * TZ;
  int t, end, stride;
  tile_range<CHUNK>(ntiles, t, end, stride);
  if (t >= end) return;
  TileDev tl = tiles[t];
  for (;;) {
    // prefetch the next tile's record; it is waited for only at the end of this iteration
    const int tn = t + stride;
    TileDev next{};
    if (tn < end) next = tiles[tn];
    // per-lane lookups go through the global table (a private copy would live in scratch memory)
    const PatchDev* __restrict__ pp = patches + tl.patch;
    const int n = tl.n;
    for (int c = threadIdx.x; c < TC; c += 256) {
      const int x = tl.x0 + c % TX, y = tl.y0 + (c / TX) % TY, z = tl.z0 + c / (TX * TY);
      if (x >= n || y >= n || z >= n) continue;
      patch_cell<T, M32, FL>(pp, words, src + tl.base, dst + tl.base, dst, n, tl.omega, x, y, z);
    }
    if (tn >= end) break;
    t = tn;
    tl = next;
  }
}

// (x, y, z) of linear cell c of an n^3 patch (x fastest); NN = n when known at compile time, else 0
template <int NN>
__device__ __forceinline__ void split_cell(int c, int n_rt, int& x, int& y, int& z) {
  const int n = NN ? NN : n_rt;
  z = c / (n * n);
  const int r = c - z * n * n;
  y = r / n;
  x = r - y * n;
}

// Linear tiles: 256 consecutive interior cells of one patch in x-fastest order, one per thread, so
// no lane idles whatever the edge (tile.x0 holds the first cell's linear index).
// FP32 state: at most 64 VGPRs, 8 waves per SIMD (the hybrid collision fits without spills);
// FP64: at least 4 waves per SIMD (its natural allocation).
template <typename T>
constexpr int patch_waves_per_simd() { return sizeof(T) == 4 ? 8 : 4; }

template <typename T, bool M32, bool FL, bool CHUNK>
__global__ __launch_bounds__(256, patch_waves_per_simd<T>()) void k_patch_step_lin(const PatchDev* __restrict__ patches,
                                                        const TileDev* __restrict__ tiles, int ntiles,
                                                        const uint32_t* __restrict__ words,
                                                        const T* __restrict__ src, T* __restrict__ dst) {
  int t, end, stride;
  tile_range<CHUNK>(ntiles, t, end, stride);
  if (t >= end) return;
  TileDev tl = tiles[t];
  for (;;) {
    const int tn = t + stride;
    TileDev next{};
    if (tn < end) next = tiles[tn];
    const PatchDev* __restrict__ pp = patches + tl.patch;
    const int n = tl.n;
    const int c = tl.x0 + (int)threadIdx.x;
    if (c < n * n * n) {
      // the edge is uniform over the workgroup: the common edges divide by constants
      int x, y, z;
      switch (n) {
        case 8: split_cell<8>(c, n, x, y, z); break;
        case 16: split_cell<16>(c, n, x, y, z); break;
        case 24: split_cell<24>(c, n, x, y, z); break;
        case 32: split_cell<32>(c, n, x, y, z); break;
        default: split_cell<0>(c, n, x, y, z); break;
      }
      patch_cell<T, M32, FL>(pp, words, src + tl.base, dst + tl.base, dst, n, tl.omega, x, y, z);
    }
    if (tn >= end) break;
    t = tn;
    tl = next;
  }
}

// Flag words from per-cell fluid bytes (patch order, x fastest): bit 31 fluid, bit q set when the
// neighbour x + c_q is not fluid.  A neighbour outside the patch lies in the patch across the crossed
// face or edge (nbr[dir_index(e)], e = the components of c_q that leave the patch), at x + c_q - n e;
// neighbouring patches share the edge n.
__global__ __launch_bounds__(256) void k_patch_words(const PatchDev* __restrict__ patches, int npatch,
                                                     const uint8_t* __restrict__ fluid, uint32_t* __restrict__ words) {
  const int p = blockIdx.y;
  if (p >= npatch) return;
  const PatchDev& P = patches[p];
  const int n = P.n;
  const int64_t cells = (int64_t)n * n * n;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (int64_t)gridDim.x * 256) {
    const int x = (int)(c % n), y = (int)((c / n) % n), z = (int)(c / ((int64_t)n * n));
    if (!fluid[P.cbase + c]) {
      words[P.cbase + c] = 0u;
      continue;
    }
    uint32_t w = FLUID;
    for (int q = 0; q < Q; q++) {
      if (q == REST) continue;
      int sx = x + CX(q), sy = y + CY(q), sz = z + CZ(q);
      const int ex = sx < 0 ? -1 : (sx >= n ? 1 : 0), ey = sy < 0 ? -1 : (sy >= n ? 1 : 0),
                ez = sz < 0 ? -1 : (sz >= n ? 1 : 0);
      int np = p;
      if (ex | ey | ez) {
        np = P.nbr[dir_index(ex, ey, ez)];
        sx -= n * ex;
        sy -= n * ey;
        sz -= n * ez;
      }
      if (!fluid[patches[np].cbase + sx + (int64_t)n * (sy + (int64_t)n * sz)]) w |= 1u << q;
    }
    words[P.cbase + c] = w;
  }
}

// Flag word of interior cell c of patch p: FLUID without bounce bits when the set has no flags.
__device__ __forceinline__ uint32_t patch_word(const uint32_t* __restrict__ words, const int64_t* __restrict__ cell_base,
                                               int p, int64_t c) {
  return words ? words[cell_base[p] + c] : FLUID;
}

// Slot that holds population q of the reference state of interior cell c = (x, y, z): the cell's own
// plane 18 - q across a bounce-back link, else the pull source (own plane or ghost record).
__device__ __forceinline__ int64_t ref_slot(int q, int n, int x, int y, int z, int64_t c, uint32_t w) {
  if ((w >> (Q - 1 - q)) & 1u) return (int64_t)(Q - 1 - q) * n * n * n + c;
  return slot_offset(q, n, x - CX(q), y - CY(q), z - CZ(q));
}

// initial state: f = feq(rho, u) per fluid interior cell, written to the slot each population is
// pulled from (own plane or own ghost record, or the own reflected plane across a bounce-back link);
// solid cells write nothing (no fluid cell reads their slots)
template <typename T>
__global__ __launch_bounds__(256) void k_patch_init(const PatchDev* __restrict__ patches, int npatch,
                                                    const int64_t* __restrict__ cell_base, const uint32_t* __restrict__ words,
                                                    const double* __restrict__ rho, const double* __restrict__ u,
                                                    T* __restrict__ s) {
  const int p = blockIdx.y;
  if (p >= npatch) return;
  const int64_t base = patches[p].base;
  const int n = patches[p].n;
  const int64_t cells = (int64_t)n * n * n;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (int64_t)gridDim.x * 256) {
    const int x = (int)(c % n), y = (int)((c / n) % n), z = (int)(c / ((int64_t)n * n));
    const uint32_t w = patch_word(words, cell_base, p, c);
    if (!(w & FLUID)) continue;
    const int64_t ui = cell_base[p] + c;
    const double r = rho ? rho[ui] : 1.0;
    const double u0 = u ? u[3 * ui] : 0.0, u1 = u ? u[3 * ui + 1] : 0.0, u2 = u ? u[3 * ui + 2] : 0.0;
    double uu = u0 * u0;
    uu += u1 * u1;
    uu += u2 * u2;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const double cu = cdot64(q, u0, u1, u2);
      const double feq = W64(q) * r *
          (1.0 + cu / (1.0 / 3.0) + (cu * cu) / (2.0 * (1.0 / 3.0) * (1.0 / 3.0)) - (uu / (2.0 * (1.0 / 3.0))));
      const T v = (sizeof(T) == 8) ? (T)feq : (T)(feq - W64(q));
      s[base + ref_slot(q, n, x, y, z, c, w)] = v;
    }
  }
}

// reference state of every interior cell (FP64), patch-major SoA blocks of 19 * n^3; 0 at solid cells
template <typename T>
__global__ __launch_bounds__(256) void k_patch_pull(const PatchDev* __restrict__ patches, int npatch,
                                                    const int64_t* __restrict__ cell_base, const uint32_t* __restrict__ words,
                                                    const T* __restrict__ s, double* __restrict__ f) {
  const int p = blockIdx.y;
  if (p >= npatch) return;
  const int64_t base = patches[p].base;
  const int n = patches[p].n;
  const int64_t cells = (int64_t)n * n * n;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (int64_t)gridDim.x * 256) {
    const int x = (int)(c % n), y = (int)((c / n) % n), z = (int)(c / ((int64_t)n * n));
    const uint32_t w = patch_word(words, cell_base, p, c);
    const bool fl = (w & FLUID) != 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      double v = 0.0;
      if (fl) {
        const T sv = s[base + ref_slot(q, n, x, y, z, c, w)];
        v = (sizeof(T) == 8) ? (double)sv : (double)sv + W64(q);
      }
      f[Q * cell_base[p] + (int64_t)q * cells + c] = v;
    }
  }
}

// rho and u of every interior cell (patch order, x fastest), in the operation order of the dense
// readout (BLB/BlockComputeCellValues.cpp:26-68: density, then momentum times the reciprocal)
// (rho = 0, u = 0 at solid cells, as the dense readout)
template <typename T>
__global__ __launch_bounds__(256) void k_patch_fields(const PatchDev* __restrict__ patches, int npatch,
                                                      const int64_t* __restrict__ cell_base, const uint32_t* __restrict__ words,
                                                      const T* __restrict__ s, double* __restrict__ rho,
                                                      double* __restrict__ u) {
  const int p = blockIdx.y;
  if (p >= npatch) return;
  const int64_t base = patches[p].base;
  const int n = patches[p].n;
  const int64_t cells = (int64_t)n * n * n;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (int64_t)gridDim.x * 256) {
    const int x = (int)(c % n), y = (int)((c / n) % n), z = (int)(c / ((int64_t)n * n));
    const uint32_t w = patch_word(words, cell_base, p, c);
    if (!(w & FLUID)) {
      const int64_t k = cell_base[p] + c;
      rho[k] = 0.0;
      u[3 * k] = u[3 * k + 1] = u[3 * k + 2] = 0.0;
      continue;
    }
    double r = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const T v = s[base + ref_slot(q, n, x, y, z, c, w)];
      const double fv = (sizeof(T) == 8) ? (double)v : (double)v + W64(q);
      r += fv;
      if (CX(q) == 1) b0 += fv;
      if (CX(q) == -1) b0 -= fv;
      if (CY(q) == 1) b1 += fv;
      if (CY(q) == -1) b1 -= fv;
      if (CZ(q) == 1) b2 += fv;
      if (CZ(q) == -1) b2 -= fv;
    }
    const double inv = 1.0 / r;
    const int64_t k = cell_base[p] + c;
    rho[k] = r;
    u[3 * k] = b0 * inv;
    u[3 * k + 1] = b1 * inv;
    u[3 * k + 2] = b2 * inv;
  }
}

// per-block partial mass / momentum deviations over interior cells, summed in fixed order later
// (fluid cells only)
template <typename T>
__global__ __launch_bounds__(256) void k_patch_moments(const PatchDev* __restrict__ patches, int npatch,
                                                       const int64_t* __restrict__ cell_base,
                                                       const uint32_t* __restrict__ words, const T* __restrict__ s,
                                                       double* __restrict__ partials) {
  const int p = blockIdx.y;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (p < npatch) {
    const int64_t base = patches[p].base;
    const int n = patches[p].n;
    const int64_t cells = (int64_t)n * n * n;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (int64_t)gridDim.x * 256) {
      const int x = (int)(c % n), y = (int)((c / n) % n), z = (int)(c / ((int64_t)n * n));
      const uint32_t w = patch_word(words, cell_base, p, c);
      if (!(w & FLUID)) continue;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const T sv = s[base + ref_slot(q, n, x, y, z, c, w)];
        const double fv = (sizeof(T) == 8) ? (double)sv : (double)sv + W64(q);
        v[0] += (sizeof(T) == 8) ? (double)sv - W64(q) : (double)sv;
        if (CX(q) == 1) v[1] += fv;
        if (CX(q) == -1) v[1] -= fv;
        if (CY(q) == 1) v[2] += fv;
        if (CY(q) == -1) v[2] -= fv;
        if (CZ(q) == 1) v[3] += fv;
        if (CZ(q) == -1) v[3] -= fv;
      }
    }
  }
  __shared__ double lds[4][4];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (threadIdx.x % 64 == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) lds[k][threadIdx.x / 64] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t slot = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) partials[4 * slot + k] = lds[k][0] + lds[k][1] + lds[k][2] + lds[k][3];
  }
}

}  // namespace lbm

using namespace lbm;

namespace {
// Candidate tilings for the autotuner: 256 threads per workgroup, 256 or 512 cells per tile.
struct Shape {
  int x, y, z;
};
// box tiles of 256 cells, and {256, 0, 0}: linear runs of 256 consecutive cells of a patch
const Shape kShapes[] = {{8, 8, 4}, {16, 8, 2}, {16, 16, 1}, {32, 8, 1}, {16, 4, 4}, {256, 0, 0}};
bool linear_tiles(const lbm_tiling& t) { return t.tile_y == 0; }

double tau_eff_of(double tau) {
  double nu = 1.0 / 3.0 * (tau - 0.5);
  nu = nu * 1.0;
  return 3.0 * nu + 0.5;
}
}  // namespace

struct lbm_patchset {
  int device = 0;
  bool fp32 = false;
  bool f32math = false;
  size_t esize = 8;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::vector<PatchDev> host;
  std::vector<int64_t> cell_base;  // interior-cell offset of each patch in user arrays
  int64_t cells = 0;               // interior cells
  int64_t elems = 0;               // stored elements per buffer
  PatchDev* d_patches = nullptr;
  int64_t* d_cell_base = nullptr;
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;
  TileDev* d_tiles = nullptr;
  int ntiles = 0;
  lbm_tiling tiling{16, 16, 1, 4, 0};
  int cus = 256;
  double* partials = nullptr;
  int64_t npartials = 0;
  double* small = nullptr;
  bool state_set = false;
  uint32_t* d_words = nullptr;     // per-cell flag words (null: every cell is fluid)
  int64_t nfluid = 0;              // fluid interior cells
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_cur = -1;  // buffer parity the graph was captured for
};

namespace {
thread_local std::string g_patch_create_error;

int pfail(lbm_patchset* ps, int code, const std::string& m) {
  if (ps) ps->err = m;
  return code;
}
#define PCHK(ps, call)                                                                                           \
  do {                                                                                                           \
    hipError_t e_ = (call);                                                                                      \
    if (e_ != hipSuccess) return pfail(ps, LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

void drop_graph(lbm_patchset* ps) {
  if (ps->gexec) (void)hipGraphExecDestroy(ps->gexec);
  if (ps->graph) (void)hipGraphDestroy(ps->graph);
  ps->gexec = nullptr;
  ps->graph = nullptr;
  ps->graph_cur = -1;
}

// tile list for a shape: largest patches first, then patch order, then z, y, x
int build_tiles(lbm_patchset* ps, const lbm_tiling& t) {
  std::vector<int> order(ps->host.size());
  for (size_t p = 0; p < order.size(); p++) order[p] = (int)p;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ps->host[a].n > ps->host[b].n; });
  std::vector<TileDev> tiles;
  for (int p : order) {
    const int n = ps->host[p].n;
    if (linear_tiles(t)) {
      const PatchDev& d = ps->host[p];
      for (int c = 0; c < n * n * n; c += t.tile_x) tiles.push_back(TileDev{d.base, d.omega, p, d.n, c, 0, 0});
      continue;
    }
    for (int z = 0; z < n; z += t.tile_z)
      for (int y = 0; y < n; y += t.tile_y)
        for (int x = 0; x < n; x += t.tile_x) {
          const PatchDev& d = ps->host[p];
          tiles.push_back(TileDev{d.base, d.omega, p, d.n, x, y, z});
        }
  }
  if (ps->d_tiles) (void)hipFree(ps->d_tiles);
  ps->d_tiles = nullptr;
  if (lbm::device_alloc(&ps->d_tiles, tiles.size() * sizeof(TileDev)) != hipSuccess)
    return pfail(ps, LBM_ERR_NOMEM, "tile table allocation failed");
  PCHK(ps, hipMemcpy(ps->d_tiles, tiles.data(), tiles.size() * sizeof(TileDev), hipMemcpyHostToDevice));
  ps->ntiles = (int)tiles.size();
  ps->tiling = t;
  drop_graph(ps);
  return LBM_OK;
}

template <typename T, bool M32, bool FL, bool CHUNK>
hipError_t launch_step_c(lbm_patchset* ps, const void* src, void* dst) {
  const lbm_tiling& t = ps->tiling;
  // blocks_per_cu = 0: one workgroup per tile (the loop body runs once)
  const int wgs = t.blocks_per_cu == 0 ? ps->ntiles : std::min(ps->ntiles, ps->cus * t.blocks_per_cu);
  const dim3 grid((unsigned)std::max(1, wgs)), block(256);
  const T* s = (const T*)src;
  T* d = (T*)dst;
#define LBM_PATCH(X, Y, Z)                                                                                     \
  if (t.tile_x == X && t.tile_y == Y && t.tile_z == Z) {                                                       \
    hipLaunchKernelGGL((k_patch_step<T, X, Y, Z, M32, FL, CHUNK>), grid, block, 0, ps->stream, ps->d_patches,   \
                       ps->d_tiles, ps->ntiles, ps->d_words, s, d);                                            \
    return hipGetLastError();                                                                                  \
  }
  LBM_PATCH(8, 8, 4)
  LBM_PATCH(16, 8, 2)
  LBM_PATCH(16, 16, 1)
  LBM_PATCH(32, 8, 1)
  LBM_PATCH(16, 4, 4)
#undef LBM_PATCH
  if (linear_tiles(t)) {
    hipLaunchKernelGGL((k_patch_step_lin<T, M32, FL, CHUNK>), grid, block, 0, ps->stream, ps->d_patches,
                       ps->d_tiles, ps->ntiles, ps->d_words, s, d);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

// one workgroup per tile runs the strided kernel (each workgroup's loop body runs once)
template <typename T, bool M32, bool FL>
hipError_t launch_step_t(lbm_patchset* ps, const void* src, void* dst) {
  return (ps->tiling.chunked && ps->tiling.blocks_per_cu) ? launch_step_c<T, M32, FL, true>(ps, src, dst)
                                                           : launch_step_c<T, M32, FL, false>(ps, src, dst);
}

template <bool FL>
hipError_t launch_step_fl(lbm_patchset* ps, int from) {
  if (ps->fp32)
    return ps->f32math ? launch_step_t<float, true, FL>(ps, ps->buf[from], ps->buf[from ^ 1])
                       : launch_step_t<float, false, FL>(ps, ps->buf[from], ps->buf[from ^ 1]);
  return launch_step_t<double, false, FL>(ps, ps->buf[from], ps->buf[from ^ 1]);
}

hipError_t launch_step(lbm_patchset* ps, int from) {
  return ps->d_words ? launch_step_fl<true>(ps, from) : launch_step_fl<false>(ps, from);
}

bool shape_supported(const lbm_tiling* t) {
  for (const Shape& s : kShapes)
    if (s.x == t->tile_x && s.y == t->tile_y && s.z == t->tile_z) return true;
  return false;
}
}  // namespace

extern "C" {

int lbm_patches_create(lbm_patchset** out, const lbm_patch* table, int npatches, int precision, int device) {
  if (!out) return LBM_ERR_ARG;
  *out = nullptr;
  lbm_patchset* ps = new lbm_patchset();
  auto bail = [&](int code, const std::string& m) {
    g_patch_create_error = m;  // lbm_patches_last_error(NULL) reports it
    lbm_patches_destroy(ps);
    return code;
  };
  if (!table || npatches < 1) return bail(LBM_ERR_ARG, "empty patch table");
  if (precision != LBM_FP64 && precision != LBM_FP32_SHIFTED && precision != LBM_FP32_SHIFTED_F32)
    return bail(LBM_ERR_ARG, "unknown precision");
  for (int p = 0; p < npatches; p++) {
    const lbm_patch& P = table[p];
    if (P.n < 2 || P.n > 256) return bail(LBM_ERR_ARG, "patch edge must be in 2..256");
    if (!(P.tau > 0.5)) return bail(LBM_ERR_ARG, "patch tau must be > 0.5");
    for (int q = 0; q < Q; q++) {
      if (q == REST) continue;
      if (P.nbr[q] < 0 || P.nbr[q] >= npatches) return bail(LBM_ERR_ARG, "neighbour index out of range");
      if (table[P.nbr[q]].n != P.n) return bail(LBM_ERR_ARG, "neighbouring patches must share their edge");
      // the neighbour across c_q must see this patch across c_{18-q}: each ghost record has one writer
      if (table[P.nbr[q]].nbr[Q - 1 - q] != p) return bail(LBM_ERR_ARG, "neighbour table must be reciprocal");
    }
  }
  ps->device = device;
  ps->fp32 = precision != LBM_FP64;
  ps->f32math = precision == LBM_FP32_SHIFTED_F32;
  ps->esize = ps->fp32 ? 4 : 8;
  int64_t base = 0, cb = 0;
  for (int p = 0; p < npatches; p++) {
    PatchDev d{};
    d.n = table[p].n;
    d.base = base;
    d.omega = 1.0 / tau_eff_of(table[p].tau);
    for (int q = 0; q < Q; q++) d.nbr[q] = q == REST ? p : table[p].nbr[q];
    d.cbase = cb;
    ps->host.push_back(d);
    ps->cell_base.push_back(cb);
    base += patch_elems(d.n);  // 64-element aligned blocks
    cb += (int64_t)d.n * d.n * d.n;
  }
  for (auto& d : ps->host)
    for (int q = 0; q < Q; q++) d.nbase[q] = ps->host[d.nbr[q]].base;
  ps->elems = base;
  ps->cells = cb;
  ps->nfluid = cb;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess)
    return bail(LBM_ERR_HIP, "device query failed");
  ps->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ps->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ps->ev0) != hipSuccess || hipEventCreate(&ps->ev1) != hipSuccess)
    return bail(LBM_ERR_HIP, "stream/event creation failed");
  const size_t bytes = (size_t)ps->elems * ps->esize;
  int maxn = 0;
  for (auto& d : ps->host) maxn = std::max(maxn, d.n);
  ps->npartials = (int64_t)npatches * 64;
  if (lbm::device_alloc(&ps->buf[0], bytes) != hipSuccess || lbm::device_alloc(&ps->buf[1], bytes) != hipSuccess ||
      lbm::device_alloc(&ps->d_patches, ps->host.size() * sizeof(PatchDev)) != hipSuccess ||
      lbm::device_alloc(&ps->d_cell_base, ps->cell_base.size() * sizeof(int64_t)) != hipSuccess ||
      lbm::device_alloc(&ps->partials, (size_t)ps->npartials * 4 * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&ps->small, 8 * sizeof(double)) != hipSuccess)
    return bail(LBM_ERR_NOMEM, "device allocation failed");
  if (hipMemcpy(ps->d_patches, ps->host.data(), ps->host.size() * sizeof(PatchDev), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ps->d_cell_base, ps->cell_base.data(), ps->cell_base.size() * sizeof(int64_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(ps->buf[0], 0, bytes) != hipSuccess || hipMemset(ps->buf[1], 0, bytes) != hipSuccess)
    return bail(LBM_ERR_HIP, "initial upload failed");
  if (build_tiles(ps, ps->tiling) != LBM_OK) return bail(LBM_ERR_HIP, ps->err);
  *out = ps;
  return LBM_OK;
}

int lbm_patches_destroy(lbm_patchset* ps) {
  if (!ps) return LBM_OK;
  if (ps->stream) (void)hipStreamSynchronize(ps->stream);
  drop_graph(ps);
  void* ptrs[] = {ps->buf[0], ps->buf[1], ps->d_patches, ps->d_cell_base, ps->d_tiles, ps->partials, ps->small,
                  ps->d_words};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (ps->ev0) (void)hipEventDestroy(ps->ev0);
  if (ps->ev1) (void)hipEventDestroy(ps->ev1);
  if (ps->stream) (void)hipStreamDestroy(ps->stream);
  delete ps;
  return LBM_OK;
}

const char* lbm_patches_last_error(const lbm_patchset* ps) {
  return ps ? ps->err.c_str() : g_patch_create_error.c_str();
}

int64_t lbm_patches_cells(const lbm_patchset* ps) { return ps ? ps->cells : -1; }

int lbm_patches_set_flags(lbm_patchset* ps, const uint8_t* fluid) {
  if (!ps) return LBM_ERR_ARG;
  PCHK(ps, hipSetDevice(ps->device));
  PCHK(ps, hipStreamSynchronize(ps->stream));
  drop_graph(ps);  // captured steps bake in the kernel variant and the word array
  if (ps->d_words) (void)hipFree(ps->d_words);
  ps->d_words = nullptr;
  ps->state_set = false;  // populations sit in pull slots that depend on the flags: init again
  ps->nfluid = ps->cells;
  if (!fluid) return LBM_OK;
  int64_t nf = 0;
  for (int64_t k = 0; k < ps->cells; k++) nf += fluid[k] != 0;
  uint8_t* d_fluid = nullptr;
  if (lbm::device_alloc(&d_fluid, (size_t)ps->cells) != hipSuccess)
    return pfail(ps, LBM_ERR_NOMEM, "flag staging allocation failed");
  if (lbm::device_alloc(&ps->d_words, (size_t)ps->cells * sizeof(uint32_t)) != hipSuccess) {
    ps->d_words = nullptr;
    (void)hipFree(d_fluid);
    return pfail(ps, LBM_ERR_NOMEM, "flag word allocation failed");
  }
  hipError_t e = hipMemcpyAsync(d_fluid, fluid, (size_t)ps->cells, hipMemcpyHostToDevice, ps->stream);
  if (e == hipSuccess) {
    int maxn = 0;
    for (auto& d : ps->host) maxn = std::max(maxn, d.n);
    const int64_t maxc = (int64_t)maxn * maxn * maxn;
    const dim3 grid((unsigned)std::min<int64_t>((maxc + 255) / 256, 64), (unsigned)ps->host.size());
    hipLaunchKernelGGL(k_patch_words, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(), d_fluid,
                       ps->d_words);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ps->stream);
  (void)hipFree(d_fluid);
  if (e != hipSuccess) {
    (void)hipFree(ps->d_words);
    ps->d_words = nullptr;
    return pfail(ps, LBM_ERR_HIP, std::string("patch flags: ") + hipGetErrorString(e));
  }
  ps->nfluid = nf;
  return LBM_OK;
}

int lbm_patches_init(lbm_patchset* ps, const double* rho, const double* u) {
  if (!ps) return LBM_ERR_ARG;
  PCHK(ps, hipSetDevice(ps->device));
  double *d_rho = nullptr, *d_u = nullptr;
  if (rho && lbm::device_alloc(&d_rho, (size_t)ps->cells * sizeof(double)) != hipSuccess)
    return pfail(ps, LBM_ERR_NOMEM, "staging allocation failed");
  if (u && lbm::device_alloc(&d_u, (size_t)ps->cells * 3 * sizeof(double)) != hipSuccess) {
    if (d_rho) (void)hipFree(d_rho);
    return pfail(ps, LBM_ERR_NOMEM, "staging allocation failed");
  }
  hipError_t e = hipSuccess;
  if (rho) e = hipMemcpy(d_rho, rho, (size_t)ps->cells * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && u) e = hipMemcpy(d_u, u, (size_t)ps->cells * 3 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    int maxn = 0;
    for (auto& d : ps->host) maxn = std::max(maxn, d.n);
    const int64_t maxc = (int64_t)maxn * maxn * maxn;
    const dim3 grid((unsigned)std::min<int64_t>((maxc + 255) / 256, 64), (unsigned)ps->host.size());
    if (ps->fp32)
      hipLaunchKernelGGL(k_patch_init<float>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                         ps->d_cell_base, ps->d_words, d_rho, d_u, (float*)ps->buf[ps->cur]);
    else
      hipLaunchKernelGGL(k_patch_init<double>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                         ps->d_cell_base, ps->d_words, d_rho, d_u, (double*)ps->buf[ps->cur]);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ps->stream);
  }
  if (d_rho) (void)hipFree(d_rho);
  if (d_u) (void)hipFree(d_u);
  if (e != hipSuccess) return pfail(ps, LBM_ERR_HIP, std::string("patch init: ") + hipGetErrorString(e));
  ps->state_set = true;
  return LBM_OK;
}

int lbm_patches_get_populations(lbm_patchset* ps, double* f) {
  if (!ps) return LBM_ERR_ARG;
  if (!f) return pfail(ps, LBM_ERR_ARG, "output is null");
  if (!ps->state_set) return pfail(ps, LBM_ERR_STATE, "state not initialised (lbm_patches_init)");
  PCHK(ps, hipSetDevice(ps->device));
  double* d_f = nullptr;
  const size_t bytes = (size_t)ps->cells * Q * sizeof(double);
  if (lbm::device_alloc(&d_f, bytes) != hipSuccess) return pfail(ps, LBM_ERR_NOMEM, "staging allocation failed");
  int maxn = 0;
  for (auto& d : ps->host) maxn = std::max(maxn, d.n);
  const int64_t maxc = (int64_t)maxn * maxn * maxn;
  const dim3 grid((unsigned)std::min<int64_t>((maxc + 255) / 256, 64), (unsigned)ps->host.size());
  if (ps->fp32)
    hipLaunchKernelGGL(k_patch_pull<float>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                       ps->d_cell_base, ps->d_words, (const float*)ps->buf[ps->cur], d_f);
  else
    hipLaunchKernelGGL(k_patch_pull<double>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                       ps->d_cell_base, ps->d_words, (const double*)ps->buf[ps->cur], d_f);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(f, d_f, bytes, hipMemcpyDeviceToHost, ps->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ps->stream);
  (void)hipFree(d_f);
  if (e != hipSuccess) return pfail(ps, LBM_ERR_HIP, std::string("patch readout: ") + hipGetErrorString(e));
  return LBM_OK;
}

int lbm_patches_get_fields(lbm_patchset* ps, double* rho, double* u) {
  if (!ps) return LBM_ERR_ARG;
  if (!rho || !u) return pfail(ps, LBM_ERR_ARG, "output is null");
  if (!ps->state_set) return pfail(ps, LBM_ERR_STATE, "state not initialised (lbm_patches_init)");
  PCHK(ps, hipSetDevice(ps->device));
  double* d_f = nullptr;
  const size_t cells = (size_t)ps->cells;
  if (lbm::device_alloc(&d_f, 4 * cells * sizeof(double)) != hipSuccess)
    return pfail(ps, LBM_ERR_NOMEM, "staging allocation failed");
  double* d_rho = d_f;
  double* d_u = d_f + cells;
  int maxn = 0;
  for (auto& d : ps->host) maxn = std::max(maxn, d.n);
  const int64_t maxc = (int64_t)maxn * maxn * maxn;
  const dim3 grid((unsigned)std::min<int64_t>((maxc + 255) / 256, 64), (unsigned)ps->host.size());
  if (ps->fp32)
    hipLaunchKernelGGL(k_patch_fields<float>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                       ps->d_cell_base, ps->d_words, (const float*)ps->buf[ps->cur], d_rho, d_u);
  else
    hipLaunchKernelGGL(k_patch_fields<double>, grid, dim3(256), 0, ps->stream, ps->d_patches, (int)ps->host.size(),
                       ps->d_cell_base, ps->d_words, (const double*)ps->buf[ps->cur], d_rho, d_u);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(rho, d_rho, cells * sizeof(double), hipMemcpyDeviceToHost, ps->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(u, d_u, 3 * cells * sizeof(double), hipMemcpyDeviceToHost, ps->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ps->stream);
  (void)hipFree(d_f);
  if (e != hipSuccess) return pfail(ps, LBM_ERR_HIP, std::string("patch fields: ") + hipGetErrorString(e));
  return LBM_OK;
}

int lbm_patches_set_tiling(lbm_patchset* ps, const lbm_tiling* t) {
  if (!ps) return LBM_ERR_ARG;
  if (!t || !shape_supported(t)) return pfail(ps, LBM_ERR_ARG, "unsupported tile shape");
  if (t->blocks_per_cu < 0 || t->blocks_per_cu > 64) return pfail(ps, LBM_ERR_ARG, "blocks_per_cu must be in 0..64");
  if (t->chunked != 0 && t->chunked != 1) return pfail(ps, LBM_ERR_ARG, "chunked must be 0 or 1");
  PCHK(ps, hipSetDevice(ps->device));
  return build_tiles(ps, *t);
}

int lbm_patches_step(lbm_patchset* ps, int nsteps, int use_graph, float* ms) {
  if (!ps) return LBM_ERR_ARG;
  if (nsteps < 0) return pfail(ps, LBM_ERR_ARG, "nsteps must be >= 0");
  if (!ps->state_set) return pfail(ps, LBM_ERR_STATE, "state not initialised (lbm_patches_init)");
  PCHK(ps, hipSetDevice(ps->device));
  PCHK(ps, hipEventRecord(ps->ev0, ps->stream));
  int s = 0;
  if (use_graph && nsteps >= 2) {
    if (!ps->gexec || ps->graph_cur != ps->cur) {
      drop_graph(ps);
      PCHK(ps, hipStreamBeginCapture(ps->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e1 = launch_step(ps, ps->cur);
      hipError_t e2 = launch_step(ps, ps->cur ^ 1);
      hipError_t e3 = hipStreamEndCapture(ps->stream, &ps->graph);
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        drop_graph(ps);
        return pfail(ps, LBM_ERR_HIP, "graph capture failed");
      }
      PCHK(ps, hipGraphInstantiate(&ps->gexec, ps->graph, nullptr, nullptr, 0));
      ps->graph_cur = ps->cur;
    }
    for (; s + 2 <= nsteps; s += 2) PCHK(ps, hipGraphLaunch(ps->gexec, ps->stream));
  }
  for (; s < nsteps; s++) {
    PCHK(ps, launch_step(ps, ps->cur));
    ps->cur ^= 1;
  }
  PCHK(ps, hipEventRecord(ps->ev1, ps->stream));
  if (ms) {
    PCHK(ps, hipEventSynchronize(ps->ev1));
    PCHK(ps, hipEventElapsedTime(ms, ps->ev0, ps->ev1));
  }
  return LBM_OK;
}

int lbm_patches_autotune(lbm_patchset* ps, lbm_tiling* chosen) {
  if (!ps) return LBM_ERR_ARG;
  if (!ps->state_set) return pfail(ps, LBM_ERR_STATE, "state not initialised (lbm_patches_init)");
  PCHK(ps, hipSetDevice(ps->device));
  // time on a copy of the state, restore afterwards
  const size_t bytes = (size_t)ps->elems * ps->esize;
  void* save = nullptr;
  if (lbm::device_alloc(&save, bytes) != hipSuccess) return pfail(ps, LBM_ERR_NOMEM, "autotune staging allocation failed");
  const int cur0 = ps->cur;
  PCHK(ps, hipMemcpy(save, ps->buf[cur0], bytes, hipMemcpyDeviceToDevice));
  lbm_tiling best{};
  float best_ms = 1e30f;
  const int bpcs[] = {4, 8, 16, 32, 64, 0};
  int rc = LBM_OK;
  for (const Shape& sh : kShapes) {
    for (int bpc : bpcs) {
      for (int chunked = 0; chunked <= (bpc ? 1 : 0) && !rc; chunked++) {
        lbm_tiling t{sh.x, sh.y, sh.z, bpc, chunked};
        rc = build_tiles(ps, t);
        if (rc) break;
        float ms = 0.f;
        rc = lbm_patches_step(ps, 2, 0, nullptr);
        if (!rc) rc = lbm_patches_step(ps, 6, 0, &ms);
        if (rc) break;
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
      if (rc) break;
    }
    if (rc) break;
  }
  hipError_t e = hipMemcpy(ps->buf[cur0], save, bytes, hipMemcpyDeviceToDevice);
  (void)hipFree(save);
  ps->cur = cur0;
  if (rc) return rc;
  if (e != hipSuccess) return pfail(ps, LBM_ERR_HIP, "autotune restore failed");
  rc = build_tiles(ps, best);
  if (rc) return rc;
  if (chosen) *chosen = best;
  return LBM_OK;
}

int lbm_patches_reduce_moments(lbm_patchset* ps, double out[4]) {
  if (!ps) return LBM_ERR_ARG;
  if (!out) return pfail(ps, LBM_ERR_ARG, "output is null");
  if (!ps->state_set) return pfail(ps, LBM_ERR_STATE, "state not initialised (lbm_patches_init)");
  PCHK(ps, hipSetDevice(ps->device));
  const int np = (int)ps->host.size();
  const dim3 grid(64, (unsigned)np);
  if (ps->fp32)
    hipLaunchKernelGGL(k_patch_moments<float>, grid, dim3(256), 0, ps->stream, ps->d_patches, np,
                       ps->d_cell_base, ps->d_words, (const float*)ps->buf[ps->cur], ps->partials);
  else
    hipLaunchKernelGGL(k_patch_moments<double>, grid, dim3(256), 0, ps->stream, ps->d_patches, np,
                       ps->d_cell_base, ps->d_words, (const double*)ps->buf[ps->cur], ps->partials);
  PCHK(ps, hipGetLastError());
  std::vector<double> h((size_t)np * 64 * 4);
  PCHK(ps, hipMemcpyAsync(h.data(), ps->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, ps->stream));
  PCHK(ps, hipStreamSynchronize(ps->stream));
  double s[4] = {0, 0, 0, 0};
  for (size_t j = 0; j < h.size() / 4; j++)
    for (int k = 0; k < 4; k++) s[k] += h[4 * j + k];
  out[0] = (double)ps->nfluid + s[0];
  for (int k = 1; k < 4; k++) out[k] = s[k];
  for (int k = 0; k < 4; k++)
    if (!std::isfinite(out[k])) return pfail(ps, LBM_ERR_NUMERIC, "non-finite population detected");
  return LBM_OK;
}

}  // extern "C"
