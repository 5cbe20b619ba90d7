// Device-side cell mapping and neighbour addressing of the dense lattice (shared by the dense
// and adjoint kernels).
#pragma once

#include "lbm_internal.h"

namespace lbm {

// Workgroup mapping.  Workgroups are dispatched round-robin over the 8 XCDs.  With g.remap the id is
// remapped so that each XCD owns one contiguous range of workgroups (a slab of planes or brick
// layers); without it all XCDs stream the same few planes at a time.  Which is faster depends on
// the lattice and the allocation (porous 256^3 FP32 rows: dispatch order +3.5 %; FP64 bricks: remap
// +1.5 %; all-fluid 512^3 rows: dispatch order 38.3k MLUPS, remap 35.5k-38.1k; 1024^3: the reverse),
// so lbm_tune times both; the untuned defaults (make_geom) are the remap for bricks and for lattices
// of 2^29 cells and more, dispatch order otherwise.
// Rows: thread (x, y) of a workgroup owns cell (64 bx + x, g.by * by + y, z).  Bricks: lane
// threadIdx.x of wavefront threadIdx.y owns cell threadIdx.x of the brick in traversal slot
// g.by * workgroup + threadIdx.y; slots run over tiles of 2 x 2 x 2 bricks (x fastest inside a tile,
// tiles in x-fastest order), which keeps a brick's z- and y-neighbours in flight together more
// often than the storage order does (FP64 +3 %; larger tiles were slower).
struct CellPos {
  int x, y, z;
  int64_t i;   // stored (bricked) index
  int64_t r;   // row-major user index x + nx * (y + ny * z)
  bool valid;
};

__device__ __forceinline__ CellPos map_cell(const DenseGeom& g) {
  CellPos c;
  const int64_t bid = blockIdx.x;
  const int64_t per = g.nblocks_padded / NXCD;
  const int64_t b = g.remap ? (bid % NXCD) * per + bid / NXCD : bid;
  if (!g.bricked) {
    const int bx = (int)(b % g.rnx);
    const int64_t r = b / g.rnx;
    const int by = (int)(r % g.rny);
    c.z = (int)(r / g.rny);
    c.x = bx * BX + threadIdx.x;
    c.y = by * g.by + threadIdx.y;
    c.valid = c.z < g.nz && c.x < g.nx && c.y < g.ny;
    c.i = c.r = (int64_t)c.x + (int64_t)g.nx * ((int64_t)c.y + (int64_t)g.ny * c.z);
    return c;
  }
  const int64_t slot = b * g.by + threadIdx.y;
  const int l = threadIdx.x;
  const int64_t tile = slot >> 3;
  const int loc = (int)(slot & 7);
  const int tx = (int)(tile % g.tnx);
  const int64_t rest = tile / g.tnx;
  const int ty = (int)(rest % g.tny);
  const int64_t tz = rest / g.tny;
  const int bx = 2 * tx + (loc & 1), by = 2 * ty + ((loc >> 1) & 1);
  const int64_t bz = 2 * tz + (loc >> 2);
  c.x = 4 * bx + (l & 3);
  c.y = 4 * by + ((l >> 2) & 3);
  c.z = (int)(4 * bz + (l >> 4));
  c.valid = bz < g.bnz && bx < g.bnx && by < g.bny && c.x < g.nx && c.y < g.ny && c.z < g.nz;
  c.i = ((int64_t)bx + (int64_t)g.bnx * ((int64_t)by + (int64_t)g.bny * bz)) * 64 + l;
  c.r = (int64_t)c.x + (int64_t)g.nx * ((int64_t)c.y + (int64_t)g.ny * c.z);
  return c;
}

// position of stored cell i
__device__ __forceinline__ CellPos pos_of(const DenseGeom& g, int64_t i) {
  CellPos c;
  c.i = i;
  c.valid = true;
  if (!g.bricked) {
    c.x = (int)(i % g.nx);
    c.y = (int)((i / g.nx) % g.ny);
    c.z = (int)(i / ((int64_t)g.nx * g.ny));
    c.r = i;
    return c;
  }
  const int64_t brick = i >> 6;
  const int l = (int)(i & 63);
  c.x = 4 * (int)(brick % g.bnx) + (l & 3);
  c.y = 4 * (int)((brick / g.bnx) % g.bny) + ((l >> 2) & 3);
  c.z = 4 * (int)(brick / ((int64_t)g.bnx * g.bny)) + (l >> 4);
  c.r = (int64_t)c.x + (int64_t)g.nx * ((int64_t)c.y + (int64_t)g.ny * c.z);
  c.valid = true;
  return c;
}

// Linear offsets of the x-1/x+1, y-1/y+1, z-1/z+1 neighbours.  Periodic axes wrap; on a
// non-periodic axis a neighbour outside the lattice is clamped to the cell itself (offset 0), so
// every address stays inside the lattice even for loads issued before the bounce bits are known
// (such a link always carries the bounce bit, and the value loaded there is replaced).
struct NbrOff {
  int64_t xm, xp, ym, yp, zm, zp;
};

// The stored index is a sum of one term per coordinate (stored_index), so a neighbour's index is
// the cell's plus the change of each term.
__device__ __forceinline__ int64_t brick_term_x(int x) { return (int64_t)((x >> 2) << 6) + (x & 3); }
__device__ __forceinline__ int64_t brick_term_y(const DenseGeom& g, int y) {
  return (int64_t)(y >> 2) * 64 * g.bnx + ((y & 3) << 2);
}
__device__ __forceinline__ int64_t brick_term_z(const DenseGeom& g, int z) {
  return (int64_t)(z >> 2) * 64 * g.bnx * g.bny + ((z & 3) << 4);
}

__device__ __forceinline__ NbrOff nbr_offsets(const DenseGeom& g, const CellPos& c) {
  NbrOff o;
  if (!g.bricked) {
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    o.xm = c.x == 0 ? (g.px ? (int64_t)(g.nx - 1) : 0) : -1;
    o.xp = c.x == g.nx - 1 ? (g.px ? -(int64_t)(g.nx - 1) : 0) : 1;
    o.ym = c.y == 0 ? (g.py ? (int64_t)(g.ny - 1) * sy : 0) : -sy;
    o.yp = c.y == g.ny - 1 ? (g.py ? -(int64_t)(g.ny - 1) * sy : 0) : sy;
    o.zm = c.z == 0 ? (g.pz ? (int64_t)(g.nz - 1) * sz : 0) : -sz;
    o.zp = c.z == g.nz - 1 ? (g.pz ? -(int64_t)(g.nz - 1) * sz : 0) : sz;
    return o;
  }
  const int64_t tx = brick_term_x(c.x), ty = brick_term_y(g, c.y), tz = brick_term_z(g, c.z);
  o.xm = c.x == 0 ? (g.px ? brick_term_x(g.nx - 1) - tx : 0) : brick_term_x(c.x - 1) - tx;
  o.xp = c.x == g.nx - 1 ? (g.px ? -tx : 0) : brick_term_x(c.x + 1) - tx;
  o.ym = c.y == 0 ? (g.py ? brick_term_y(g, g.ny - 1) - ty : 0) : brick_term_y(g, c.y - 1) - ty;
  o.yp = c.y == g.ny - 1 ? (g.py ? -ty : 0) : brick_term_y(g, c.y + 1) - ty;
  o.zm = c.z == 0 ? (g.pz ? brick_term_z(g, g.nz - 1) - tz : 0) : brick_term_z(g, c.z - 1) - tz;
  o.zp = c.z == g.nz - 1 ? (g.pz ? -tz : 0) : brick_term_z(g, c.z + 1) - tz;
  return o;
}

// Neighbour offsets in the address space of interleaved populations (aos = 1).  There the address
// of population q of the cell at (x, y, z) is a sum of one term per coordinate plus 64 q whenever a
// 64-cell group is a brick (bricks) or a row segment of whole groups (rows with nx a multiple of 64,
// the only interleaved rows make_geom allows): bricks: x (x >> 2) 1216 + (x & 3), y (y >> 2) 1216 bnx
// + 4 (y & 3), z (z >> 2) 1216 bnx bny + 16 (z & 3); rows: x (x >> 6) 1216 + (x & 63), y 1216 y nx/64,
// z 1216 z ny nx/64 (1216 = 19 x 64).  So a neighbour's address is the cell's plus per-axis deltas,
// exactly as a neighbour's index is in the plane layout, and the kernels pay no more per access.
__device__ __forceinline__ int64_t aos_term_x(const DenseGeom& g, int x) {
  return g.bricked ? (int64_t)(x >> 2) * (Q * 64) + (x & 3) : (int64_t)(x >> 6) * (Q * 64) + (x & 63);
}
__device__ __forceinline__ int64_t aos_term_y(const DenseGeom& g, int y) {
  return g.bricked ? (int64_t)(y >> 2) * (Q * 64) * g.bnx + ((y & 3) << 2) : (int64_t)y * (Q * 64) * (g.nx >> 6);
}
__device__ __forceinline__ int64_t aos_term_z(const DenseGeom& g, int z) {
  return g.bricked ? (int64_t)(z >> 2) * (Q * 64) * g.bnx * g.bny + ((z & 3) << 4)
                   : (int64_t)z * (Q * 64) * (g.nx >> 6) * g.ny;
}
__device__ __forceinline__ NbrOff aos_offsets(const DenseGeom& g, const CellPos& c) {
  NbrOff o;
  const int64_t tx = aos_term_x(g, c.x), ty = aos_term_y(g, c.y), tz = aos_term_z(g, c.z);
  o.xm = c.x == 0 ? (g.px ? aos_term_x(g, g.nx - 1) - tx : 0) : aos_term_x(g, c.x - 1) - tx;
  o.xp = c.x == g.nx - 1 ? (g.px ? -tx + aos_term_x(g, 0) : 0) : aos_term_x(g, c.x + 1) - tx;
  o.ym = c.y == 0 ? (g.py ? aos_term_y(g, g.ny - 1) - ty : 0) : aos_term_y(g, c.y - 1) - ty;
  o.yp = c.y == g.ny - 1 ? (g.py ? -ty + aos_term_y(g, 0) : 0) : aos_term_y(g, c.y + 1) - ty;
  o.zm = c.z == 0 ? (g.pz ? aos_term_z(g, g.nz - 1) - tz : 0) : aos_term_z(g, c.z - 1) - tz;
  o.zp = c.z == g.nz - 1 ? (g.pz ? -tz + aos_term_z(g, 0) : 0) : aos_term_z(g, c.z + 1) - tz;
  return o;
}

// offset of the cell x - c_q (the source of population q in a pull)
__device__ __forceinline__ int64_t src_off(int q, const NbrOff& o) {
  int64_t d = 0;
  if (CX(q) == 1) d += o.xm;
  if (CX(q) == -1) d += o.xp;
  if (CY(q) == 1) d += o.ym;
  if (CY(q) == -1) d += o.yp;
  if (CZ(q) == 1) d += o.zm;
  if (CZ(q) == -1) d += o.zp;
  return d;
}
// offset of the cell x + c_q
__device__ __forceinline__ int64_t dst_off(int q, const NbrOff& o) {
  int64_t d = 0;
  if (CX(q) == 1) d += o.xp;
  if (CX(q) == -1) d += o.xm;
  if (CY(q) == 1) d += o.yp;
  if (CY(q) == -1) d += o.ym;
  if (CZ(q) == 1) d += o.zp;
  if (CZ(q) == -1) d += o.zm;
  return d;
}

// Addresses of one cell's populations and of its neighbours' in the hot kernels: own(q) is the
// cell's population q, own(q) + src_off(q, o) / dst_off(q, o) that of the neighbour x - c_q / x + c_q.
// Planes: base = i, o = index offsets, own(q) = q N + i.  Interleaved: base = the address of the
// cell's population 0, o = aos_offsets, own(q) = base + 64 q.
template <bool AOS>
struct CellAddr {
  int64_t base, N;
  NbrOff o;
  __device__ __forceinline__ int64_t own(int q) const { return AOS ? base + 64 * q : (int64_t)q * N + base; }
  __device__ __forceinline__ int64_t pull(int q) const { return own(q) + src_off(q, o); }
  __device__ __forceinline__ int64_t push(int q) const { return own(q) + dst_off(q, o); }
};
template <bool AOS>
__device__ __forceinline__ CellAddr<AOS> cell_addr(const DenseGeom& g, const CellPos& c) {
  CellAddr<AOS> a;
  a.N = g.N;
  if (AOS) {
    a.base = pop_at<true>(g.N, 0, c.i);
    a.o = aos_offsets(g, c);
  } else {
    a.base = c.i;
    a.o = nbr_offsets(g, c);
  }
  return a;
}

// index of the stored value that becomes population q of cell i in the reference state
template <bool AOS>
__device__ __forceinline__ int64_t pull_index(int q, uint32_t w, int64_t i, int64_t N, const NbrOff& o) {
  const bool bounce = (w >> (Q - 1 - q)) & 1u;  // x - c_q = x + c_{18-q} is not fluid
  return bounce ? pop_at<AOS>(N, Q - 1 - q, i) : pop_at<AOS>(N, q, i + src_off(q, o));
}
__device__ __forceinline__ int64_t pull_index(const DenseGeom& g, int q, uint32_t w, int64_t i, const NbrOff& o) {
  return g.aos ? pull_index<true>(q, w, i, g.N, o) : pull_index<false>(q, w, i, g.N, o);
}

// the same for AA's SWAPPED layout (post-collision values reflected in the own slots)
template <bool AOS>
__device__ __forceinline__ int64_t swapped_index(int q, uint32_t w, int64_t i, int64_t N, const NbrOff& o) {
  const bool bounce = (w >> (Q - 1 - q)) & 1u;
  return bounce ? pop_at<AOS>(N, q, i) : pop_at<AOS>(N, Q - 1 - q, i + src_off(q, o));
}

// index of population q of the reference state at cell i in the layout g.layout
template <bool AOS>
__device__ __forceinline__ int64_t state_index(const DenseGeom& g, int q, uint32_t w, int64_t i, const NbrOff& o) {
  if (g.layout == LAYOUT_AA_NATURAL) return pop_at<AOS>(g.N, q, i);
  if (g.layout == LAYOUT_AA_SWAPPED) return swapped_index<AOS>(q, w, i, g.N, o);
  return pull_index<AOS>(q, w, i, g.N, o);
}
__device__ __forceinline__ int64_t state_index(const DenseGeom& g, int q, uint32_t w, int64_t i, const NbrOff& o) {
  return g.aos ? state_index<true>(g, q, w, i, o) : state_index<false>(g, q, w, i, o);
}

}  // namespace lbm
