// Device-side cell mapping and neighbour addressing of the dense lattice (shared by the dense
// and adjoint kernels).
#pragma once

#include "lbm_internal.h"

namespace lbm {

// XCD-aware block mapping.  Workgroups are dispatched round-robin over the 8 XCDs, so the
// linear id is remapped to give each XCD one contiguous range of blocks (a stack of z-planes):
// the y- and z-neighbour rows a block pulls from were then mostly fetched into the same L2.
struct CellPos {
  int x, y, z;
  int64_t i;
  bool valid;
};

__device__ __forceinline__ CellPos map_cell(const DenseGeom& g) {
  CellPos c;
  const int64_t bid = blockIdx.x;
  const int64_t per = g.nblocks_padded / NXCD;
  const int64_t b = (bid % NXCD) * per + bid / NXCD;
  const int bx = (int)(b % g.nbx);
  const int64_t r = b / g.nbx;
  const int by = (int)(r % g.nby);
  c.z = (int)(r / g.nby);
  c.x = bx * BX + threadIdx.x;
  c.y = by * BY + threadIdx.y;
  c.valid = (b < g.nblocks) && c.x < g.nx && c.y < g.ny;
  c.i = (int64_t)c.x + (int64_t)g.nx * ((int64_t)c.y + (int64_t)g.ny * c.z);
  return c;
}

// Linear offsets of the x-1/x+1, y-1/y+1, z-1/z+1 neighbours.  Periodic axes wrap; on a
// non-periodic axis a neighbour outside the lattice is clamped to the cell itself (offset 0), so
// every address stays inside the lattice even for loads issued before the bounce bits are known
// (such a link always carries the bounce bit, and the value loaded there is replaced).
struct NbrOff {
  int64_t xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ NbrOff nbr_offsets(const DenseGeom& g, const CellPos& c) {
  NbrOff o;
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  o.xm = c.x == 0 ? (g.px ? (int64_t)(g.nx - 1) : 0) : -1;
  o.xp = c.x == g.nx - 1 ? (g.px ? -(int64_t)(g.nx - 1) : 0) : 1;
  o.ym = c.y == 0 ? (g.py ? (int64_t)(g.ny - 1) * sy : 0) : -sy;
  o.yp = c.y == g.ny - 1 ? (g.py ? -(int64_t)(g.ny - 1) * sy : 0) : sy;
  o.zm = c.z == 0 ? (g.pz ? (int64_t)(g.nz - 1) * sz : 0) : -sz;
  o.zp = c.z == g.nz - 1 ? (g.pz ? -(int64_t)(g.nz - 1) * sz : 0) : sz;
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

// index of the stored value that becomes population q of cell i in the reference state
__device__ __forceinline__ int64_t pull_index(int q, uint32_t w, int64_t i, int64_t N, const NbrOff& o) {
  const bool bounce = (w >> (Q - 1 - q)) & 1u;  // x - c_q = x + c_{18-q} is not fluid
  return bounce ? (int64_t)(Q - 1 - q) * N + i : (int64_t)q * N + i + src_off(q, o);
}

// the same for AA's SWAPPED layout (post-collision values reflected in the own slots)
__device__ __forceinline__ int64_t swapped_index(int q, uint32_t w, int64_t i, int64_t N, const NbrOff& o) {
  const bool bounce = (w >> (Q - 1 - q)) & 1u;
  return bounce ? (int64_t)q * N + i : (int64_t)(Q - 1 - q) * N + i + src_off(q, o);
}

// index of population q of the reference state at cell i in the layout g.layout
__device__ __forceinline__ int64_t state_index(const DenseGeom& g, int q, uint32_t w, int64_t i, const NbrOff& o) {
  if (g.layout == LAYOUT_AA_NATURAL) return (int64_t)q * g.N + i;
  if (g.layout == LAYOUT_AA_SWAPPED) return swapped_index(q, w, i, g.N, o);
  return pull_index(q, w, i, g.N, o);
}

}  // namespace lbm
