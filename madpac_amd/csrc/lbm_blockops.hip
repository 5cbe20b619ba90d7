// The block solver's per-block and per-Peano-cell operators on the device block store.
//
// Reference counterpart: the kernels that RegularGrid2RegularBlockSolver
// (BLB/mappings/RegularGrid2RegularBlockSolver.cpp) calls from its events, one block or one Peano
// cell at a time: computeDensity / computeVelocity (BLB/BlockComputeCellValues.cpp:26-68), the BGK
// collision (BLB/collisionmodels/CollisionModelWrapper.cpp:28-39, BlockBGK.cpp:18-45), the boundary
// conditions after the collision (BLB/boundaryconditions/BlockBoundaryConditionManager.cpp:47-62),
// and the swap streaming of BLB/nonlocaltreatment/AABlockStreaming.cpp (streamingLocal :19-32,
// streamingNonLocal :37-53, swapPdfsInsideCells :80-92).  Here each operator is one launch over a
// list of blocks (or Peano cells), one workgroup per block (or cell).
//
// Per-block operators stage the block's 4,104 populations (32.8 KB, AoS as the store keeps them)
// through LDS with coalesced 16-byte loads and stores, run the chosen operators there in the
// reference's order, and write the block back once.  (Persistent workgroups that load the next
// block into registers while one is in LDS measured 7 % slower at 256^3: they need 128 VGPRs.)
// Fused, touchVertexFirstTime's four kernels move each block across HBM once instead of four
// times.  The cross-block swap (lbm_blocks_stream_nonlocal) touches only the boundary-crossing
// populations and works in place in HBM: its swaps are disjoint (each cross-block link belongs to
// exactly one Peano cell and one entry of its table), so all cells of a call run concurrently.
// Whole iterations (lbm_blocks_iterate) take one launch each: the block pass writes its outgoing
// crossing values into an exchange array and the next pass applies the swaps from there, before
// the in-cell swap and the collision.  All of it is HBM-bound byte movement plus the FP64
// collision, computed in the reference's operation order (bit-identical; the build uses
// -ffp-contract=off).
#include <algorithm>
#include <string>
#include <vector>

#include "lbm_blockstore.h"

namespace lbm {

// ---------------------------------------------------------------------------------------------
// CellTopology::getNonLocalStreamingIndex (BLB/nonlocaltreatment/CellTopology.cpp:435-465, even
// block size :317-418), built at compile time.  Stated on the Peano cell's 6^3 sub-cells: vertex v
// owns the octant [3 v_d, 3 v_d + 3) and sub-cell s of it is cell s_d + 3 - 6 v_d of v's block; for
// vertex pairs i < j, sub-cells s of i's octant (z, y, x order) and directions q, the entry
// (i, j, cell of s, cell of s + c_q in block j, q) exists when s + c_q lies on j's side of the
// cell's centre planes along every axis.  Packed i | j << 3 | c1 << 6 | c2 << 14 | q << 22.
constexpr int NL_ENTRIES = 504;
struct NonLocalTable {
  uint32_t e[NL_ENTRIES];
  int n;
};

constexpr NonLocalTable make_nonlocal_table() {
  NonLocalTable t{};
  int n = 0;
  for (int i = 0; i < 8; i++)
    for (int j = i + 1; j < 8; j++) {
      const int vi[3] = {i & 1, (i >> 1) & 1, i >> 2}, vj[3] = {j & 1, (j >> 1) & 1, j >> 2};
      for (int sz = 3 * vi[2]; sz < 3 * vi[2] + 3; sz++)
        for (int sy = 3 * vi[1]; sy < 3 * vi[1] + 3; sy++)
          for (int sx = 3 * vi[0]; sx < 3 * vi[0] + 3; sx++)
            for (int q = 0; q < Q; q++) {
              const int t3[3] = {sx + CX(q), sy + CY(q), sz + CZ(q)};
              bool side = true;
              for (int d = 0; d < 3; d++) side = side && ((t3[d] >= 3 ? 1 : 0) == vj[d]);
              if (!side) continue;
              const int c1 = (sx + 3 - 6 * vi[0]) + BS * ((sy + 3 - 6 * vi[1]) + BS * (sz + 3 - 6 * vi[2]));
              const int c2 = (t3[0] + 3 - 6 * vj[0]) + BS * ((t3[1] + 3 - 6 * vj[1]) + BS * (t3[2] + 3 - 6 * vj[2]));
              if (n < NL_ENTRIES)
                t.e[n] = (uint32_t)i | (uint32_t)j << 3 | (uint32_t)c1 << 6 | (uint32_t)c2 << 14 | (uint32_t)q << 22;
              n++;
            }
    }
  t.n = n;
  return t;
}
static_assert(make_nonlocal_table().n == NL_ENTRIES, "D3Q19 on 6^3 blocks has 504 cross-block links per Peano cell");

__constant__ NonLocalTable kNonLocal = make_nonlocal_table();

// ---------------------------------------------------------------------------------------------
// A block's boundary-crossing populations: the 1,008 entries (cell, q) whose target cell + c_q lies
// in a neighbour block, across one of 18 faces and edges (a D3Q19 link never crosses a corner).
// Entries are grouped by direction; the segment of a direction lists its entries in (cell, q)
// order, and the segment of the opposite direction lists their partners (cell + c_q wrapped into
// the neighbour, 18 - q) in the same order, so entry k of a segment pairs with entry k of the
// neighbour's opposite segment.  Packed cell | q << 8 | direction << 13.
constexpr int NX_ENTRIES = 1008, NDIR = 18;
struct CrossTable {
  uint32_t e[NX_ENTRIES];
  int start[NDIR], len[NDIR], opp[NDIR];
  int16_t entry_of[BP];  // crossing entry of (cell, q), or -1
  int n;
};

// direction index of each neighbour offset (ox + 1) + 3 (oy + 1) + 9 (oz + 1) with one or two
// nonzero components, numbered in z, y, x order; -1 for the centre and the corners
struct DirMap {
  int d[27];
};
constexpr DirMap make_dir_map() {
  DirMap m{};
  int n = 0;
  for (int k = 0; k < 27; k++) {
    const int x = k % 3 - 1, y = (k / 3) % 3 - 1, z = k / 9 - 1;
    const int nz = (x != 0) + (y != 0) + (z != 0);
    m.d[k] = (nz >= 1 && nz <= 2) ? n++ : -1;
  }
  return m;
}

constexpr CrossTable make_cross_table() {
  CrossTable t{};
  const DirMap dm = make_dir_map();
  int off27[NDIR] = {};
  for (int k = 0; k < 27; k++)
    if (dm.d[k] >= 0) off27[dm.d[k]] = k;
  for (int d = 0; d < NDIR; d++) t.opp[d] = dm.d[26 - off27[d]];
  // direction of every (cell, q), and the partner slot across it
  int16_t dir_of[BP] = {}, partner_of[BP] = {};
  for (int c = 0; c < BC; c++)
    for (int q = 0; q < Q; q++) {
      const int x = c % BS + CX(q), y = (c / BS) % BS + CY(q), z = c / (BS * BS) + CZ(q);
      const int ox = x < 0 ? -1 : (x >= BS ? 1 : 0), oy = y < 0 ? -1 : (y >= BS ? 1 : 0),
                oz = z < 0 ? -1 : (z >= BS ? 1 : 0);
      dir_of[c * Q + q] = (int16_t)dm.d[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
      partner_of[c * Q + q] =
          (int16_t)(((x - BS * ox) + BS * ((y - BS * oy) + BS * (z - BS * oz))) * Q + (Q - 1 - q));
    }
  for (int k = 0; k < BP; k++) t.entry_of[k] = -1;
  int n = 0;
  for (int d = 0; d < NDIR; d++) {
    t.start[d] = n;
    const bool first = d < t.opp[d];  // the lower index of a pair lists (cell, q) in order
    const int src = first ? d : t.opp[d];
    for (int k = 0; k < BP; k++) {
      if (dir_of[k] != src) continue;
      const int slot = first ? k : partner_of[k];  // the partner in the neighbour across src
      if (n < NX_ENTRIES) {
        t.e[n] = (uint32_t)(slot / Q) | (uint32_t)(slot % Q) << 8 | (uint32_t)d << 13;
        t.entry_of[slot] = (int16_t)n;
      }
      n++;
    }
    t.len[d] = n - t.start[d];
  }
  t.n = n;
  return t;
}
static_assert(make_cross_table().n == NX_ENTRIES, "a 6^3 block has 1,008 boundary-crossing populations");

// The two tables describe the same links: every cross-block entry (i, j, c1, c2, q) of a Peano cell
// pairs crossing entry (c1, q) of block i, in the direction from vertex i to vertex j, with its
// partner (c2, 18 - q) in the opposite segment of block j; and partnering is an involution.
constexpr bool cross_tables_agree() {
  const CrossTable t = make_cross_table();
  const NonLocalTable nl = make_nonlocal_table();
  const DirMap dm = make_dir_map();
  for (int d = 0; d < NDIR; d++)
    if (t.len[d] != t.len[t.opp[d]] || t.opp[t.opp[d]] != d) return false;
  for (int e = 0; e < NX_ENTRIES; e++) {
    const int d = (t.e[e] >> 13) & 31, p = t.start[t.opp[d]] + (e - t.start[d]);
    if (t.entry_of[(t.e[e] & 255) * Q + ((t.e[e] >> 8) & 31)] != e) return false;
    if (t.start[t.opp[t.opp[d]]] + (p - t.start[t.opp[d]]) != e) return false;
  }
  for (int k = 0; k < NL_ENTRIES; k++) {
    const uint32_t E = nl.e[k];
    const int i = E & 7, j = (E >> 3) & 7, c1 = (E >> 6) & 255, c2 = (E >> 14) & 255, q = (E >> 22) & 31;
    const int o = ((j & 1) - (i & 1) + 1) + 3 * (((j >> 1) & 1) - ((i >> 1) & 1) + 1) + 9 * ((j >> 2) - (i >> 2) + 1);
    const int e1 = t.entry_of[c1 * Q + q], e2 = t.entry_of[c2 * Q + (Q - 1 - q)];
    if (e1 < 0 || e2 < 0) return false;
    const int d = (t.e[e1] >> 13) & 31;
    if (d != dm.d[o] || t.start[t.opp[d]] + (e1 - t.start[d]) != e2) return false;
  }
  return true;
}
static_assert(cross_tables_agree(), "crossing entries and the Peano cell's cross-block table disagree");

__constant__ CrossTable kCross = make_cross_table();

// The segment bounds as compile-time constants: the direction of crossing entry e and its partner
// entry in the neighbour (e + start[opp d] - start[d]) follow from comparisons alone, with no table
// load in front of the neighbour lookup.
struct CrossSegments {
  int start[NDIR], delta[NDIR];
};
constexpr CrossSegments make_cross_segments() {
  const CrossTable t = make_cross_table();
  CrossSegments s{};
  for (int d = 0; d < NDIR; d++) {
    s.start[d] = t.start[d];
    s.delta[d] = t.start[t.opp[d]] - t.start[d];
  }
  return s;
}
constexpr CrossSegments kSeg = make_cross_segments();

__host__ __device__ constexpr void cross_segment(int e, int& d, int& partner) {
  d = 0;
  int delta = kSeg.delta[0];
#pragma unroll
  for (int k = 1; k < NDIR; k++)
    if (e >= kSeg.start[k]) {
      d = k;
      delta = kSeg.delta[k];
    }
  partner = e + delta;
}

// the comparisons give every entry the direction the table packs into it, and the partner entry
// that the table's segments pair with it
constexpr bool cross_segments_agree() {
  const CrossTable t = make_cross_table();
  for (int e = 0; e < NX_ENTRIES; e++) {
    int d = -1, pe = -1;
    cross_segment(e, d, pe);
    if (d != (int)((t.e[e] >> 13) & 31u) || pe != t.start[t.opp[d]] + (e - t.start[d])) return false;
  }
  return true;
}
static_assert(cross_segments_agree(), "segment bounds and the crossing table disagree");

// ---------------------------------------------------------------------------------------------
// OP_XIN: the previous iteration's cross-block swaps, each valid entry taking its partner's
// outgoing value from the exchange array; OP_XOUT: this iteration's outgoing values into it.
enum : unsigned { OP_SWAP_IN = 1u, OP_VALUES = 2u, OP_COLLIDE = 4u, OP_BOUNDARY = 8u, OP_LOCAL = 16u,
                  OP_XIN = 32u, OP_XOUT = 64u };

struct BlockOpArgs {
  const int64_t* list;  // block ids, one workgroup each
  int64_t n;
  double* pdf;
  double* density;
  double* velocity;
  const uint8_t* inner;
  const uint8_t* mask;       // overlap cells (hasMinusOneEntries), or null
  const BlockLinkDev* links;  // grouped by list position
  const int32_t* off;         // links of list position p: off[p] .. off[p + 1] - 1
  double omega;               // 1.0 / tau
  const double* xin;          // exchange array read by OP_XIN (1,008 values per block)
  double* xout;               // exchange array written by OP_XOUT
  const uint32_t* valid;      // per block 32 words: bit e = entry e is swapped (listed, both cells inner)
  const int64_t* nbr;         // per block 18 neighbour blocks by direction (-1: none)
};

// one boundary link after the collision on the outgoing slot f (see lbm_blocks_treat_boundary); rho
// and (u0, u1, u2) are the cell's stored density and velocity
__device__ __forceinline__ void block_boundary_f64(double& f, const BlockLinkDev& L, double rho, double u0,
                                                   double u1, double u2) {
  const int q = Q - 1 - L.q_in;
  if (L.type == 1 || L.type == 5) {
    f -= moving_wall_f64(q, L.u, rho);
    return;
  }
  if (L.type < 2 || L.type > 4) return;
  double rt, t0, t1, t2;
  if (L.type == 2) {
    rt = L.rho; t0 = u0; t1 = u1; t2 = u2;
  } else {
    rt = L.type == 3 ? rho : L.rho;
    t0 = L.u[0]; t1 = L.u[1]; t2 = L.u[2];
  }
  double uu = t0 * t0;
  uu += t1 * t1;
  uu += t2 * t2;
  const double fin = feq_f64(L.q_in, rt, t0, t1, t2, uu);
  f = L.type == 4 ? fin : -f + fin + feq_f64(q, rt, t0, t1, t2, uu);
}

template <unsigned OPS>
__global__ __launch_bounds__(256) void k_block_ops(BlockOpArgs a) {
  constexpr bool WRITES = (OPS & (OP_SWAP_IN | OP_COLLIDE | OP_BOUNDARY | OP_LOCAL | OP_XIN | OP_XOUT)) != 0;
  constexpr bool MOMENTS = (OPS & (OP_VALUES | OP_COLLIDE | OP_BOUNDARY)) != 0;
  __shared__ __align__(16) double s[BP];
  __shared__ double srho[MOMENTS ? BC : 1];
  __shared__ double su[MOMENTS ? BV : 1];
  __shared__ uint8_t sin[BC];
  const int t = threadIdx.x;
  const int64_t p = blockIdx.x;
  if (p >= a.n) return;
  const int64_t b = a.list[p];
  // this block's boundary links (offsets read at the start, not between the collision's phases)
  const int32_t link0 = (OPS & OP_BOUNDARY) ? a.off[p] : 0, link1 = (OPS & OP_BOUNDARY) ? a.off[p + 1] : 0;
  // the previous iteration's cross-block swaps: each thread fetches the partner values of its (at
  // most 4) swapped crossing entries from the exchange array while the block's own loads are in
  // flight, and puts them into the block once it is in LDS.  The swap-mask words and neighbour ids
  // are loaded first and together (an entry's direction needs no load), then the block's pieces,
  // then the partner values: two memory round trips in front of the block's LDS writes, with the
  // block's loads in flight during both.
  constexpr int NXT = (NX_ENTRIES + 255) / 256;
  double xv[NXT];
  int xs[NXT];    // OP_XIN: the block slot a partner value goes to, -1 when not swapped
  int slot[NXT];  // the block slot of each entry (cell * 19 + q)
  uint32_t vw[NXT];
  int64_t nb[NXT];
  int pe[NXT];
  if (OPS & (OP_XIN | OP_XOUT)) {
#pragma unroll
    for (int i = 0; i < NXT; i++) {
      const int e = min(t + 256 * i, NX_ENTRIES - 1);
      if (OPS & OP_XIN) {
        int d;
        cross_segment(e, d, pe[i]);
        vw[i] = a.valid[b * 32 + (e >> 5)];
        nb[i] = a.nbr[b * NDIR + d];
      }
      const uint32_t E = kCross.e[e];
      slot[i] = (int)(E & 255u) * Q + (int)((E >> 8) & 31u);
    }
  }
  const BlockPieces own = load_block(a.pdf + b * BP, t);
  const uint8_t in_t = t < BC ? a.inner[b * BC + t] : 0;
  // the cell's stored density (kept by skipped cells), its velocity when the values are not
  // recomputed, and its overlap flag, loaded here so that no load waits between the collision's
  // phases
  const int tb = min(t, BC - 1);
  double rho_in = 0.0, u_in0 = 0.0, u_in1 = 0.0, u_in2 = 0.0;
  uint8_t skip_in = 0;
  if (MOMENTS) {
    rho_in = a.density[b * BC + tb];
    if (!(OPS & OP_VALUES)) {
      const double* ui = a.velocity + b * BV + 3 * tb;
      u_in0 = ui[0];
      u_in1 = ui[1];
      u_in2 = ui[2];
    }
    if (a.mask) skip_in = a.mask[b * BC + tb];
  }
  if (OPS & OP_XIN) {
#pragma unroll
    for (int i = 0; i < NXT; i++) {
      const int e = t + 256 * i;
      // a set bit means a listed Peano cell swaps the entry, so the neighbour exists
      const bool swapped = e < NX_ENTRIES && ((vw[i] >> (e & 31)) & 1u);
      xv[i] = swapped ? a.xin[nb[i] * NX_ENTRIES + pe[i]] : 0.0;
      xs[i] = swapped ? slot[i] : -1;
    }
  }
  put_block(own, s, t);
  if (t < BC) sin[t] = in_t;
  __syncthreads();
  if (OPS & OP_XIN) {
#pragma unroll
    for (int i = 0; i < NXT; i++)
      if (xs[i] >= 0) s[xs[i]] = xv[i];
    __syncthreads();
  }
  {
    const int c = t;
    if (c < BC) {
      double* f = s + c * Q;
      if (OPS & OP_SWAP_IN) {
        // the previous iteration's swapPdfsInsideCells (every cell, inner or not)
#pragma unroll
        for (int q = 0; q < Q / 2; q++) {
          const double v = f[q];
          f[q] = f[Q - 1 - q];
          f[Q - 1 - q] = v;
        }
      }
      if (MOMENTS) {
        const bool act = sin[c] && !skip_in;
        double rho, u0, u1, u2;
        if (OPS & OP_VALUES) {
          // computeDensity (:36-39) and computeVelocity (:57-61); skipped cells keep their density
          // and get u = 0
          rho = rho_in;
          u0 = u1 = u2 = 0.0;
          if (act) {
            rho = 0.0;
#pragma unroll
            for (int q = 0; q < Q; q++) rho += f[q];
            double b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
            for (int q = 0; q < Q; q++) {
              if (CX(q) == 1) b0 += f[q];
              if (CX(q) == -1) b0 -= f[q];
              if (CY(q) == 1) b1 += f[q];
              if (CY(q) == -1) b1 -= f[q];
              if (CZ(q) == 1) b2 += f[q];
              if (CZ(q) == -1) b2 -= f[q];
            }
            const double inv = 1.0 / rho;
            u0 = b0 * inv;
            u1 = b1 * inv;
            u2 = b2 * inv;
            a.density[b * BC + c] = rho;
          }
          double* uo = a.velocity + b * BV + 3 * c;
          uo[0] = u0;
          uo[1] = u1;
          uo[2] = u2;
        } else {
          rho = rho_in;
          u0 = u_in0;
          u1 = u_in1;
          u2 = u_in2;
        }
        srho[c] = rho;
        su[3 * c] = u0;
        su[3 * c + 1] = u1;
        su[3 * c + 2] = u2;
        if ((OPS & OP_COLLIDE) && act) {
          // computeFeq (BLB/BlockComputeCellValues.cpp:14-20) and BlockBGK (BlockBGK.cpp:43)
          double uu = u0 * u0;
          uu += u1 * u1;
          uu += u2 * u2;
#pragma unroll
          for (int q = 0; q < Q; q++) f[q] -= a.omega * (f[q] - feq_f64(q, rho, u0, u1, u2, uu));
        }
      }
    }
    if (OPS & OP_BOUNDARY) {
      __syncthreads();
      for (int32_t k = link0 + t; k < link1; k += 256) {
        const BlockLinkDev L = a.links[k];
        block_boundary_f64(s[L.cell * Q + Q - 1 - L.q_in], L, srho[L.cell], su[3 * L.cell], su[3 * L.cell + 1],
                           su[3 * L.cell + 2]);
      }
    }
    if (OPS & OP_LOCAL) {
      // streamingLocal: the link (c, q < 9) <-> (c + c_q, 18 - q) is swapped by thread c when both
      // cells are inner; every slot belongs to one link, so the swaps are disjoint
      __syncthreads();
      if (c < BC && sin[c]) {
        const int x = c % BS, y = (c / BS) % BS, z = c / (BS * BS);
#pragma unroll
        for (int q = 0; q < Q / 2; q++) {
          const int nx = x + CX(q), ny = y + CY(q), nz = z + CZ(q);
          if (nx < 0 || nx >= BS || ny < 0 || ny >= BS || nz < 0 || nz >= BS) continue;
          const int nb = nx + BS * (ny + BS * nz);
          if (!sin[nb]) continue;
          const double v = s[c * Q + q];
          s[c * Q + q] = s[nb * Q + Q - 1 - q];
          s[nb * Q + Q - 1 - q] = v;
        }
      }
    }
    if (WRITES) {
      __syncthreads();
      store_block(s, a.pdf + b * BP, t);
      if (OPS & OP_XOUT) {
        double* xo = a.xout + b * NX_ENTRIES;
#pragma unroll
        for (int i = 0; i < NXT; i++)
          if (t + 256 * i < NX_ENTRIES) xo[t + 256 * i] = s[slot[i]];
      }
    }
  }
}

// treatBoundaryAfterCollision for a flat list of links, one thread each, in place in HBM
__global__ __launch_bounds__(256) void k_block_boundary(const BlockLinkDev* __restrict__ links, int64_t n,
                                                        double* __restrict__ pdf, const double* __restrict__ density,
                                                        const double* __restrict__ velocity) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const BlockLinkDev L = links[k];
  const double* u = velocity + L.block * BV + 3 * L.cell;
  block_boundary_f64(pdf[L.block * BP + L.cell * Q + Q - 1 - L.q_in], L, density[L.block * BC + L.cell], u[0], u[1],
                     u[2]);
}

// streamingNonLocal: one workgroup per Peano cell, one thread per table entry
__global__ __launch_bounds__(512) void k_block_nonlocal(const int64_t* __restrict__ cells, int64_t n,
                                                        double* __restrict__ pdf, const uint8_t* __restrict__ inner) {
  const int64_t p = blockIdx.x;
  const int e = threadIdx.x;
  if (p >= n || e >= NL_ENTRIES) return;
  const uint32_t E = kNonLocal.e[e];
  const int i = E & 7u, j = (E >> 3) & 7u, c1 = (E >> 6) & 255u, c2 = (E >> 14) & 255u, q = (E >> 22) & 31u;
  const int64_t bi = cells[8 * p + i], bj = cells[8 * p + j];
  if (!inner[bi * BC + c1] || !inner[bj * BC + c2]) return;
  double* x = pdf + bi * BP + c1 * Q + q;
  double* y = pdf + bj * BP + c2 * Q + (Q - 1 - q);
  const double v = *x;
  *x = *y;
  *y = v;
}

// Which Peano cell swaps each crossing entry: every entry belongs to exactly one entry
// (i, j, c1, c2, q) of a Peano cell's cross-block table, from block i's side as (c1, q) or from
// block j's side as (c2, 18 - q).  owner[e] is the vertex position the block has in that cell, and
// other[e] the partner's cell (c2 or c1), both fixed by the table.
struct EntryOwner {
  uint8_t owner[NX_ENTRIES];
  uint8_t other[NX_ENTRIES];
  bool ok;
};
constexpr EntryOwner make_entry_owner() {
  EntryOwner t{};
  const NonLocalTable nl = make_nonlocal_table();
  const CrossTable x = make_cross_table();
  bool seen[NX_ENTRIES] = {};
  t.ok = true;
  for (int k = 0; k < NL_ENTRIES; k++) {
    const uint32_t E = nl.e[k];
    const int i = E & 7, j = (E >> 3) & 7, c1 = (E >> 6) & 255, c2 = (E >> 14) & 255, q = (E >> 22) & 31;
    const int e1 = x.entry_of[c1 * Q + q], e2 = x.entry_of[c2 * Q + (Q - 1 - q)];
    if (e1 < 0 || e2 < 0 || seen[e1] || seen[e2]) t.ok = false;
    if (e1 >= 0 && e2 >= 0) {
      seen[e1] = seen[e2] = true;
      t.owner[e1] = (uint8_t)i;
      t.other[e1] = (uint8_t)c2;
      t.owner[e2] = (uint8_t)j;
      t.other[e2] = (uint8_t)c1;
    }
  }
  for (int e = 0; e < NX_ENTRIES; e++) t.ok = t.ok && seen[e];
  return t;
}
static_assert(make_entry_owner().ok, "every crossing entry belongs to exactly one cross-block link of a Peano cell");
__constant__ EntryOwner kOwner = make_entry_owner();

// The swap mask of lbm_blocks_iterate: bit e of block b is set when crossing entry e is swapped by a
// listed Peano cell, i.e. b sits at vertex position owner[e] of a listed cell (bit owner[e] of
// vpos[b]) and both cells of the link are inner.  One workgroup per block writes all 32 words of
// its mask (blocks in no listed cell get zeros), so no clearing pass and no atomics are needed.
__global__ __launch_bounds__(256) void k_block_valid(int64_t nblocks, const uint8_t* __restrict__ vpos,
                                                     const int64_t* __restrict__ nbr,
                                                     const uint8_t* __restrict__ inner, uint32_t* __restrict__ valid) {
  const int64_t b = blockIdx.x;
  if (b >= nblocks) return;
  const int t = threadIdx.x;
  const uint32_t vp = vpos[b];
#pragma unroll
  for (int k = 0; k < (NX_ENTRIES + 255) / 256; k++) {
    const int e = t + 256 * k;
    bool on = false;
    if (e < NX_ENTRIES && ((vp >> kOwner.owner[e]) & 1u)) {
      const uint32_t E = kCross.e[e];
      const int64_t nb = nbr[b * NDIR + ((E >> 13) & 31u)];
      on = nb >= 0 && inner[b * BC + (E & 255u)] && inner[nb * BC + kOwner.other[e]];
    }
    const unsigned long long m = __ballot(on);
    // lanes 0-31 of a wavefront hold word e >> 5 for even (t >> 5), lanes 32-63 the next word
    if ((t & 31) == 0) valid[b * 32 + (e >> 5)] = (uint32_t)(m >> (t & 32));
  }
}

// ---------------------------------------------------------------------------------------------
// The vertex-overlap exchange between ranks.  parallel::BlockBoundaryIndex4Pdfs::initBoundaryIndex
// (BLB/parallel/BlockBoundaryIndex4Pdfs.cpp:30-77): cells in z, y, x order, q in order, every
// population whose target cell + c_q lies outside the block, with its incoming slot 19 cell + 18 - q
// and outgoing slot 19 cell + q.  BlockBoundaryIndex (BLB/BlockBoundaryIndex.cpp:15-88): the cells
// with a coordinate below 2 or above 3, in increasing index.
constexpr int PB_ENTRIES = 1008, BB_CELLS = 208;
struct PdfBoundaryTable {
  int16_t in[PB_ENTRIES], out[PB_ENTRIES];
  int n;
};
constexpr PdfBoundaryTable make_pdf_boundary_table() {
  PdfBoundaryTable t{};
  int n = 0;
  for (int c = 0; c < BC; c++)
    for (int q = 0; q < Q; q++) {
      const int x = c % BS + CX(q), y = (c / BS) % BS + CY(q), z = c / (BS * BS) + CZ(q);
      if (x >= 0 && x < BS && y >= 0 && y < BS && z >= 0 && z < BS) continue;
      if (n < PB_ENTRIES) {
        t.in[n] = (int16_t)(c * Q + Q - 1 - q);
        t.out[n] = (int16_t)(c * Q + q);
      }
      n++;
    }
  t.n = n;
  return t;
}
static_assert(make_pdf_boundary_table().n == PB_ENTRIES, "a 6^3 block has 1,008 populations leaving it");

// the populations leaving a block are its crossing entries, each listed once
constexpr bool pdf_boundary_is_crossing_set() {
  const PdfBoundaryTable p = make_pdf_boundary_table();
  const CrossTable t = make_cross_table();
  bool seen[NX_ENTRIES] = {};
  for (int i = 0; i < PB_ENTRIES; i++) {
    const int e = t.entry_of[p.out[i]];
    if (e < 0 || seen[e]) return false;
    seen[e] = true;
  }
  return true;
}
static_assert(pdf_boundary_is_crossing_set(), "leaving populations and crossing entries disagree");

struct BoundaryCellTable {
  int16_t c[BB_CELLS];
  int n;
};
constexpr BoundaryCellTable make_boundary_cell_table() {
  BoundaryCellTable t{};
  int n = 0;
  for (int c = 0; c < BC; c++) {
    const int x = c % BS, y = (c / BS) % BS, z = c / (BS * BS);
    if (!(x < 2 || x > BS - 3 || y < 2 || y > BS - 3 || z < 2 || z > BS - 3)) continue;
    if (n < BB_CELLS) t.c[n] = (int16_t)c;
    n++;
  }
  t.n = n;
  return t;
}
static_assert(make_boundary_cell_table().n == BB_CELLS, "a 6^3 block has 208 cells near its boundary");

__constant__ PdfBoundaryTable kPdfBoundary = make_pdf_boundary_table();
__constant__ BoundaryCellTable kBoundaryCells = make_boundary_cell_table();

enum : int { PD_INIT = 0, PD_COMPUTE = 1, PD_MERGE = 2 };

// initPdfDiff / computePdfDiff / mergePdfDiff (BLB/RegularGridBlockVertex.cpp:96-141) on the listed
// blocks, one workgroup per block and up to four entries per thread (their loads issued together)
template <int MODE>
__global__ __launch_bounds__(256) void k_block_pdf_diff(const int64_t* __restrict__ list, int64_t n,
                                                        double* __restrict__ pdf, double* __restrict__ diff,
                                                        const double* __restrict__ neighbour) {
  const int64_t p = blockIdx.x;
  if (p >= n) return;
  const int64_t b = list[p];
  const int t = threadIdx.x;
  constexpr int NT = (PB_ENTRIES + 255) / 256;
  double* f = pdf + b * BP;
  double* dv = diff + b * PB_ENTRIES;
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const int i = t + 256 * k;
    if (i >= PB_ENTRIES) break;
    if (MODE == PD_INIT) {
      dv[i] = -f[kPdfBoundary.out[i]];
    } else if (MODE == PD_COMPUTE) {
      dv[i] = dv[i] + f[kPdfBoundary.in[i]];
    } else if (dv[i] == 0.0) {
      const int s = kPdfBoundary.in[i];
      f[s] = f[s] + neighbour[p * PB_ENTRIES + i];
    }
  }
}

enum : int { BD_RESET = 0, BD_SET = 1, BD_MERGE = 2 };

// resetDensityOnBlockBoundary / setDensityOnBlockBoundary / mergeDensityOnBoundary (:51-77,
// :144-160): block-boundary cell i of each listed block, where it is not inner but on the geometry
// boundary
template <int MODE>
__global__ __launch_bounds__(256) void k_block_boundary_density(const int64_t* __restrict__ list, int64_t n,
                                                                double* __restrict__ density,
                                                                const uint8_t* __restrict__ inner,
                                                                const uint8_t* __restrict__ boundary,
                                                                double* __restrict__ bdens,
                                                                const double* __restrict__ neighbour) {
  const int64_t p = blockIdx.x;
  const int i = threadIdx.x;
  if (p >= n || i >= BB_CELLS) return;
  const int64_t b = list[p];
  const int c = kBoundaryCells.c[i];
  const bool on = !inner[b * BC + c] && boundary[b * BC + c];
  if (MODE == BD_RESET)
    bdens[b * BB_CELLS + i] = 0.0;
  else if (MODE == BD_SET)
    bdens[b * BB_CELLS + i] = on ? density[b * BC + c] : 0.0;
  else if (on)
    density[b * BC + c] += neighbour[p * BB_CELLS + i];
}

// rows of a per-block array for the listed blocks, in list order (width values per block)
__global__ __launch_bounds__(256) void k_block_rows(const int64_t* __restrict__ list, int64_t n,
                                                    const double* __restrict__ src, int width,
                                                    double* __restrict__ dst) {
  const int64_t p = blockIdx.x;
  if (p >= n) return;
  const int64_t b = list[p];
  for (int i = threadIdx.x; i < width; i += 256) dst[p * width + i] = src[b * width + i];
}

}  // namespace lbm

using namespace lbm;

namespace {

// a block list: ids in range, each block once
int upload_blocks(lbm_blockstore* bs, const int64_t* blocks, int64_t n, int64_t** dst, int64_t* cap) {
  if (n < 0 || (n > 0 && !blocks)) return bfail(bs, LBM_ERR_ARG, "block list is null");
  if (n > INT32_MAX) return bfail(bs, LBM_ERR_ARG, "too many blocks in one call");
  std::vector<uint8_t> seen((size_t)bs->nblocks, 0);
  for (int64_t k = 0; k < n; k++) {
    const int64_t b = blocks[k];
    if (b < 0 || b >= bs->nblocks) return bfail(bs, LBM_ERR_ARG, "block " + std::to_string(k) + ": id out of range");
    if (seen[(size_t)b]) return bfail(bs, LBM_ERR_ARG, "block " + std::to_string(b) + " appears twice in the list");
    seen[(size_t)b] = 1;
  }
  if (int rc = ensure_capacity(bs, dst, cap, std::max<int64_t>(n, 1), "block list")) return rc;
  if (n) BCHK(bs, hipMemcpyAsync(*dst, blocks, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, bs->stream));
  return LBM_OK;
}

// a Peano cell list: 8 distinct blocks per cell, each (block, vertex position) once per call
int upload_cells(lbm_blockstore* bs, const int64_t* cells, int64_t n, int64_t** dst, int64_t* cap) {
  if (n < 0 || (n > 0 && !cells)) return bfail(bs, LBM_ERR_ARG, "cell list is null");
  if (n > INT32_MAX / 8) return bfail(bs, LBM_ERR_ARG, "too many cells in one call");
  std::vector<uint8_t> used((size_t)bs->nblocks, 0);
  for (int64_t k = 0; k < n; k++) {
    const int64_t* v = cells + 8 * k;
    for (int i = 0; i < 8; i++) {
      if (v[i] < 0 || v[i] >= bs->nblocks) return bfail(bs, LBM_ERR_ARG, "cell " + std::to_string(k) + ": block id out of range");
      for (int j = 0; j < i; j++)
        if (v[j] == v[i]) return bfail(bs, LBM_ERR_ARG, "cell " + std::to_string(k) + ": a block appears at two vertices");
      uint8_t& u = used[(size_t)v[i]];
      if (u & (1u << i))
        return bfail(bs, LBM_ERR_ARG, "block " + std::to_string(v[i]) + " is vertex " + std::to_string(i) + " of two cells");
      u |= (uint8_t)(1u << i);
    }
  }
  if (int rc = ensure_capacity(bs, dst, cap, std::max<int64_t>(8 * n, 1), "cell list")) return rc;
  if (n) BCHK(bs, hipMemcpyAsync(*dst, cells, (size_t)n * 8 * sizeof(int64_t), hipMemcpyHostToDevice, bs->stream));
  return LBM_OK;
}

// boundary links: fields in range, each (block, cell, q_in) once; with a block list (n_list >= 0)
// every link's block is listed and the links are grouped by list position (offsets uploaded)
int upload_boundary(lbm_blockstore* bs, const lbm_block_boundary* links, int64_t n, const int64_t* list,
                    int64_t n_list) {
  if (n < 0 || (n > 0 && !links)) return bfail(bs, LBM_ERR_ARG, "link list is null");
  if (n > INT32_MAX) return bfail(bs, LBM_ERR_ARG, "too many links in one call");
  std::vector<int64_t> keys((size_t)n);
  for (int64_t k = 0; k < n; k++) {
    const lbm_block_boundary& L = links[k];
    const std::string at = "link " + std::to_string(k) + ": ";
    if (L.block < 0 || L.block >= bs->nblocks) return bfail(bs, LBM_ERR_ARG, at + "block id out of range");
    if (L.cell < 0 || L.cell >= BC) return bfail(bs, LBM_ERR_ARG, at + "cell must be in 0..215");
    if (L.q_in < 0 || L.q_in >= Q) return bfail(bs, LBM_ERR_ARG, at + "direction must be in 0..18");
    if (L.type < 0 || L.type > 5) return bfail(bs, LBM_ERR_ARG, at + "type must be a BoundaryType in 0..5");
    keys[(size_t)k] = (L.block * BC + L.cell) * Q + L.q_in;
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
    return bfail(bs, LBM_ERR_ARG, "two links on the same (block, cell, direction)");
  std::vector<lbm_block_boundary> grouped;
  const lbm_block_boundary* src = links;
  if (n_list >= 0) {
    std::vector<int32_t> pos((size_t)bs->nblocks, -1);
    for (int64_t k = 0; k < n_list; k++) pos[(size_t)list[k]] = (int32_t)k;
    std::vector<int32_t> off((size_t)n_list + 1, 0);
    for (int64_t k = 0; k < n; k++) {
      const int32_t p = pos[(size_t)links[k].block];
      if (p < 0) return bfail(bs, LBM_ERR_ARG, "link " + std::to_string(k) + ": its block is not in the block list");
      off[(size_t)p + 1]++;
    }
    for (int64_t k = 0; k < n_list; k++) off[(size_t)k + 1] += off[(size_t)k];
    grouped.resize((size_t)n);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < n; k++) grouped[(size_t)fill[(size_t)pos[(size_t)links[k].block]]++] = links[k];
    src = grouped.data();
    if (int rc = ensure_capacity(bs, &bs->d_off, &bs->cap_off, n_list + 1, "link offsets")) return rc;
    BCHK(bs, hipMemcpyAsync(bs->d_off, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, bs->stream));
    // the copy reads off after this function returns: wait before it goes out of scope
    BCHK(bs, hipStreamSynchronize(bs->stream));
  }
  if (int rc = ensure_capacity(bs, &bs->d_bnd, &bs->cap_bnd, std::max<int64_t>(n, 1), "link table")) return rc;
  if (n) {
    BCHK(bs, hipMemcpyAsync(bs->d_bnd, src, (size_t)n * sizeof(BlockLinkDev), hipMemcpyHostToDevice, bs->stream));
    BCHK(bs, hipStreamSynchronize(bs->stream));
  }
  return LBM_OK;
}

BlockOpArgs op_args(lbm_blockstore* bs, int64_t n, double tau, int use_overlap) {
  BlockOpArgs a{};
  a.list = bs->d_list;
  a.n = n;
  a.pdf = bs->pdf;
  a.density = bs->density;
  a.velocity = bs->velocity;
  a.inner = bs->inner;
  a.mask = use_overlap ? bs->mask : nullptr;
  a.links = bs->d_bnd;
  a.off = bs->d_off;
  a.omega = 1.0 / tau;
  return a;
}

template <unsigned OPS>
int launch_ops(lbm_blockstore* bs, const BlockOpArgs& a) {
  if (a.n) {
    hipLaunchKernelGGL(k_block_ops<OPS>, dim3((unsigned)a.n), dim3(256), 0, bs->stream, a);
    BCHK(bs, hipGetLastError());
  }
  return LBM_OK;
}

int launch_nonlocal(lbm_blockstore* bs, const int64_t* d_cells, int64_t n) {
  if (n) {
    hipLaunchKernelGGL(k_block_nonlocal, dim3((unsigned)n), dim3(512), 0, bs->stream, d_cells, n, bs->pdf, bs->inner);
    BCHK(bs, hipGetLastError());
  }
  return LBM_OK;
}

int begin(lbm_blockstore* bs) {
  BCHK(bs, hipEventRecord(bs->ev0, bs->stream));
  return LBM_OK;
}

int finish(lbm_blockstore* bs) {
  BCHK(bs, hipEventRecord(bs->ev1, bs->stream));
  BCHK(bs, hipEventSynchronize(bs->ev1));
  BCHK(bs, hipEventElapsedTime(&bs->last_ms, bs->ev0, bs->ev1));
  return LBM_OK;
}

int check_tau(lbm_blockstore* bs, double tau) {
  if (!(tau > 0.0) || !(tau < 1e300)) return bfail(bs, LBM_ERR_ARG, "tau must be positive and finite");
  return LBM_OK;
}

template <unsigned OPS>
int run_block_op(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double tau, int use_overlap) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (OPS & OP_COLLIDE)
    if (int rc = check_tau(bs, tau)) return rc;
  if (int rc = upload_blocks(bs, blocks, n, &bs->d_list, &bs->cap_list)) return rc;
  if (int rc = begin(bs)) return rc;
  if (int rc = launch_ops<OPS>(bs, op_args(bs, n, tau, use_overlap))) return rc;
  return finish(bs);
}

// The exchange buffers, allocated and cleared on first use; their contents persist between calls.
int ensure_exchange(lbm_blockstore* bs) {
  if (bs->d_pdfdiff) return LBM_OK;
  const size_t nb = (size_t)bs->nblocks;
  double* d = nullptr;
  double* r = nullptr;
  uint8_t* f = nullptr;
  if (device_alloc(&d, nb * PB_ENTRIES * sizeof(double)) != hipSuccess ||
      device_alloc(&r, nb * BB_CELLS * sizeof(double)) != hipSuccess || device_alloc(&f, nb * BC) != hipSuccess) {
    if (d) (void)hipFree(d);
    if (r) (void)hipFree(r);
    return bfail(bs, LBM_ERR_NOMEM, "exchange buffer allocation failed");
  }
  bs->d_pdfdiff = d;
  bs->d_bdens = r;
  bs->boundary = f;
  BCHK(bs, hipMemsetAsync(d, 0, nb * PB_ENTRIES * sizeof(double), bs->stream));
  BCHK(bs, hipMemsetAsync(r, 0, nb * BB_CELLS * sizeof(double), bs->stream));
  BCHK(bs, hipMemsetAsync(f, 0, nb * BC, bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

// Validates the list, prepares the buffers and, for a merge, stages the neighbour's values (width
// per listed block; host or device memory).
int exchange_prologue(lbm_blockstore* bs, const int64_t* blocks, int64_t n, const double* in, int width,
                      const char* what) {
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = ensure_exchange(bs)) return rc;
  if (int rc = upload_blocks(bs, blocks, n, &bs->d_list, &bs->cap_list)) return rc;
  if (width) {
    if (n && !in) return bfail(bs, LBM_ERR_ARG, std::string(what) + " are null");
    if (int rc = ensure_capacity(bs, &bs->d_xfer, &bs->cap_xfer, std::max<int64_t>(n * width, 1), "exchange staging"))
      return rc;
    if (n) BCHK(bs, hipMemcpyAsync(bs->d_xfer, in, (size_t)(n * width) * sizeof(double), hipMemcpyDefault, bs->stream));
  }
  return LBM_OK;
}

template <int MODE>
int run_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n, const double* neighbour) {
  if (!bs) return LBM_ERR_ARG;
  const int width = MODE == PD_MERGE ? PB_ENTRIES : 0;
  if (int rc = exchange_prologue(bs, blocks, n, neighbour, width, "neighbour differences")) return rc;
  if (int rc = begin(bs)) return rc;
  if (n) {
    hipLaunchKernelGGL(k_block_pdf_diff<MODE>, dim3((unsigned)n), dim3(256), 0, bs->stream, bs->d_list, n, bs->pdf,
                       bs->d_pdfdiff, bs->d_xfer);
    BCHK(bs, hipGetLastError());
  }
  return finish(bs);
}

template <int MODE>
int run_boundary_density(lbm_blockstore* bs, const int64_t* blocks, int64_t n, const double* neighbour) {
  if (!bs) return LBM_ERR_ARG;
  const int width = MODE == BD_MERGE ? BB_CELLS : 0;
  if (int rc = exchange_prologue(bs, blocks, n, neighbour, width, "neighbour boundary densities")) return rc;
  if (int rc = begin(bs)) return rc;
  if (n) {
    hipLaunchKernelGGL(k_block_boundary_density<MODE>, dim3((unsigned)n), dim3(256), 0, bs->stream, bs->d_list, n,
                       bs->density, bs->inner, bs->boundary, bs->d_bdens, bs->d_xfer);
    BCHK(bs, hipGetLastError());
  }
  return finish(bs);
}

// rows of an exchange buffer for the listed blocks into out (host or device memory)
int get_rows(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double* out, bool pdf_diff) {
  if (!bs) return LBM_ERR_ARG;
  if (n && !out) return bfail(bs, LBM_ERR_ARG, "output is null");
  if (int rc = exchange_prologue(bs, blocks, n, nullptr, 0, "")) return rc;
  const int width = pdf_diff ? PB_ENTRIES : BB_CELLS;
  if (int rc = ensure_capacity(bs, &bs->d_xfer, &bs->cap_xfer, std::max<int64_t>(n * width, 1), "exchange staging"))
    return rc;
  if (n) {
    hipLaunchKernelGGL(k_block_rows, dim3((unsigned)n), dim3(256), 0, bs->stream, bs->d_list, n,
                       pdf_diff ? bs->d_pdfdiff : bs->d_bdens, width, bs->d_xfer);
    BCHK(bs, hipGetLastError());
    BCHK(bs, hipMemcpyAsync(out, bs->d_xfer, (size_t)(n * width) * sizeof(double), hipMemcpyDefault, bs->stream));
  }
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_blocks_set_inner(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* inner) {
  if (!bs) return LBM_ERR_ARG;
  if (!inner) return bfail(bs, LBM_ERR_ARG, "inner is null");
  if (int rc = check_range(bs, first, count)) return rc;
  BCHK(bs, hipSetDevice(bs->device));
  std::vector<uint8_t> m((size_t)count * BC);
  for (size_t k = 0; k < m.size(); k++) m[k] = inner[k] ? 1 : 0;
  BCHK(bs, hipMemcpyAsync(bs->inner + first * BC, m.data(), m.size(), hipMemcpyHostToDevice, bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  bs->inner_gen++;
  return LBM_OK;
}

int lbm_blocks_compute_cell_values(lbm_blockstore* bs, const int64_t* blocks, int64_t n, int use_overlap) {
  return run_block_op<OP_VALUES>(bs, blocks, n, 1.0, use_overlap);
}

int lbm_blocks_collide(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double tau, int use_overlap) {
  return run_block_op<OP_COLLIDE>(bs, blocks, n, tau, use_overlap);
}

int lbm_blocks_stream_local(lbm_blockstore* bs, const int64_t* blocks, int64_t n) {
  return run_block_op<OP_LOCAL>(bs, blocks, n, 1.0, 0);
}

int lbm_blocks_swap_inside(lbm_blockstore* bs, const int64_t* blocks, int64_t n) {
  return run_block_op<OP_SWAP_IN>(bs, blocks, n, 1.0, 0);
}

int lbm_blocks_treat_boundary(lbm_blockstore* bs, const lbm_block_boundary* links, int64_t n) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = upload_boundary(bs, links, n, nullptr, -1)) return rc;
  if (int rc = begin(bs)) return rc;
  if (n) {
    hipLaunchKernelGGL(k_block_boundary, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, bs->stream, bs->d_bnd, n,
                       bs->pdf, bs->density, bs->velocity);
    BCHK(bs, hipGetLastError());
  }
  return finish(bs);
}

int lbm_blocks_stream_nonlocal(lbm_blockstore* bs, const int64_t* cells, int64_t n) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = upload_cells(bs, cells, n, &bs->d_list, &bs->cap_list)) return rc;
  if (int rc = begin(bs)) return rc;
  if (int rc = launch_nonlocal(bs, bs->d_list, n)) return rc;
  return finish(bs);
}

int lbm_blocks_vertex_first(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double tau, int use_overlap,
                            const lbm_block_boundary* links, int64_t nlinks) {
  if (!bs) return LBM_ERR_ARG;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = check_tau(bs, tau)) return rc;
  if (int rc = upload_blocks(bs, blocks, n, &bs->d_list, &bs->cap_list)) return rc;
  if (int rc = upload_boundary(bs, links, nlinks, blocks, n)) return rc;
  if (int rc = begin(bs)) return rc;
  if (int rc = launch_ops<OP_VALUES | OP_COLLIDE | OP_BOUNDARY | OP_LOCAL>(bs, op_args(bs, n, tau, use_overlap))) return rc;
  return finish(bs);
}

// One launch per iteration: iteration t's handleCell swaps and touchVertexLastTime swap run at the
// start of iteration t + 1's block pass.  Each block pass writes its blocks' outgoing crossing
// values (after the local streaming) into an exchange array, and the next pass takes each swapped
// entry's value from the partner's outgoing values there (two arrays, alternating), which is what
// the cross-block swap of the reference leaves in that slot; the blocks' own slots keep their
// values where no swap happens.  A last pass applies the final iteration's swaps.
int lbm_blocks_iterate(lbm_blockstore* bs, const int64_t* blocks, int64_t nblocks_list, const int64_t* cells,
                       int64_t ncells, double tau, const lbm_block_boundary* links, int64_t nlinks, int nsteps) {
  if (!bs) return LBM_ERR_ARG;
  if (nsteps < 0) return bfail(bs, LBM_ERR_ARG, "nsteps must be >= 0");
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = check_tau(bs, tau)) return rc;
  if (int rc = upload_blocks(bs, blocks, nblocks_list, &bs->d_list, &bs->cap_list)) return rc;
  if (int rc = upload_cells(bs, cells, ncells, &bs->d_cells, &bs->cap_cells)) return rc;
  if (int rc = upload_boundary(bs, links, nlinks, blocks, nblocks_list)) return rc;
  // the blocks of a cell must be listed blocks
  std::vector<uint8_t> listed((size_t)bs->nblocks, 0);
  for (int64_t k = 0; k < nblocks_list; k++) listed[(size_t)blocks[k]] = 1;
  for (int64_t k = 0; k < 8 * ncells; k++)
    if (!listed[(size_t)cells[k]])
      return bfail(bs, LBM_ERR_ARG, "cell " + std::to_string(k / 8) + ": block " + std::to_string(cells[k]) +
                                        " is not in the block list");
  if (int rc = ensure_capacity(bs, &bs->d_xch, &bs->cap_xch, 2 * bs->nblocks * NX_ENTRIES, "exchange arrays")) return rc;
  const bool same = bs->iter_inner_gen == bs->inner_gen && bs->iter_cells.size() == (size_t)(8 * ncells) &&
                    std::equal(bs->iter_cells.begin(), bs->iter_cells.end(), cells);
  if (!same) {
    // the neighbour of each block across each face and edge, as the cells give it (they must agree),
    // and the swap mask; kept for the next call with the same cells and inner flags
    bs->iter_cells.clear();
    bs->iter_inner_gen = ~0ull;
    constexpr DirMap dm = make_dir_map();
    std::vector<int64_t> nbr((size_t)bs->nblocks * NDIR, -1);
    std::vector<uint8_t> vpos((size_t)bs->nblocks, 0);  // bit i: the block is vertex i of a listed cell
    for (int64_t p = 0; p < ncells; p++) {
      const int64_t* v = cells + 8 * p;
      for (int i = 0; i < 8; i++) vpos[(size_t)v[i]] |= (uint8_t)(1u << i);
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
          const int o = ((j & 1) - (i & 1) + 1) + 3 * (((j >> 1) & 1) - ((i >> 1) & 1) + 1) + 9 * ((j >> 2) - (i >> 2) + 1);
          if (dm.d[o] < 0) continue;
          int64_t& slot = nbr[(size_t)v[i] * NDIR + dm.d[o]];
          if (slot >= 0 && slot != v[j])
            return bfail(bs, LBM_ERR_ARG, "cells disagree on the neighbour of block " + std::to_string(v[i]));
          slot = v[j];
        }
    }
    if (int rc = ensure_capacity(bs, &bs->d_nbr, &bs->cap_nbr, bs->nblocks * NDIR, "neighbour table")) return rc;
    if (int rc = ensure_capacity(bs, &bs->d_valid, &bs->cap_valid, bs->nblocks * 32, "swap mask")) return rc;
    if (int rc = ensure_capacity(bs, &bs->d_vpos, &bs->cap_vpos, bs->nblocks, "vertex positions")) return rc;
    BCHK(bs, hipMemcpyAsync(bs->d_nbr, nbr.data(), nbr.size() * sizeof(int64_t), hipMemcpyHostToDevice, bs->stream));
    BCHK(bs, hipMemcpyAsync(bs->d_vpos, vpos.data(), vpos.size(), hipMemcpyHostToDevice, bs->stream));
    hipLaunchKernelGGL(k_block_valid, dim3((unsigned)bs->nblocks), dim3(256), 0, bs->stream, bs->nblocks, bs->d_vpos,
                       bs->d_nbr, bs->inner, bs->d_valid);
    BCHK(bs, hipGetLastError());
    BCHK(bs, hipStreamSynchronize(bs->stream));  // nbr and vpos are host memory of this call
    bs->iter_cells.assign(cells, cells + 8 * ncells);
    bs->iter_inner_gen = bs->inner_gen;
  }
  BlockOpArgs a = op_args(bs, nblocks_list, tau, 0);
  a.valid = bs->d_valid;
  a.nbr = bs->d_nbr;
  double* x[2] = {bs->d_xch, bs->d_xch + bs->nblocks * NX_ENTRIES};
  if (int rc = begin(bs)) return rc;
  for (int s = 0; s < nsteps; s++) {
    a.xin = x[(s + 1) & 1];
    a.xout = x[s & 1];
    const int rc = s == 0 ? launch_ops<OP_VALUES | OP_COLLIDE | OP_BOUNDARY | OP_LOCAL | OP_XOUT>(bs, a)
                          : launch_ops<OP_XIN | OP_SWAP_IN | OP_VALUES | OP_COLLIDE | OP_BOUNDARY | OP_LOCAL | OP_XOUT>(bs, a);
    if (rc) return rc;
  }
  if (nsteps > 0) {  // the last iteration's handleCell and touchVertexLastTime
    a.xin = x[(nsteps - 1) & 1];
    if (int rc = launch_ops<OP_XIN | OP_SWAP_IN>(bs, a)) return rc;
  }
  return finish(bs);
}

int lbm_blocks_set_boundary(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* boundary) {
  if (!bs) return LBM_ERR_ARG;
  if (!boundary) return bfail(bs, LBM_ERR_ARG, "boundary flags are null");
  if (int rc = check_range(bs, first, count)) return rc;
  BCHK(bs, hipSetDevice(bs->device));
  if (int rc = ensure_exchange(bs)) return rc;
  std::vector<uint8_t> m((size_t)count * BC);
  for (size_t k = 0; k < m.size(); k++) m[k] = boundary[k] ? 1 : 0;
  BCHK(bs, hipMemcpyAsync(bs->boundary + first * BC, m.data(), m.size(), hipMemcpyHostToDevice, bs->stream));
  BCHK(bs, hipStreamSynchronize(bs->stream));
  return LBM_OK;
}

int lbm_blocks_init_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n) {
  return run_pdf_diff<PD_INIT>(bs, blocks, n, nullptr);
}

int lbm_blocks_compute_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n) {
  return run_pdf_diff<PD_COMPUTE>(bs, blocks, n, nullptr);
}

int lbm_blocks_get_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double* out) {
  return get_rows(bs, blocks, n, out, true);
}

int lbm_blocks_merge_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n, const double* neighbour) {
  return run_pdf_diff<PD_MERGE>(bs, blocks, n, neighbour);
}

int lbm_blocks_reset_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n) {
  return run_boundary_density<BD_RESET>(bs, blocks, n, nullptr);
}

int lbm_blocks_set_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n, int dynamic_geometry) {
  // without a dynamic geometry the reference returns before touching the buffer (:57-59); the list
  // is still checked
  if (!dynamic_geometry) {
    if (!bs) return LBM_ERR_ARG;
    BCHK(bs, hipSetDevice(bs->device));
    return upload_blocks(bs, blocks, n, &bs->d_list, &bs->cap_list);
  }
  return run_boundary_density<BD_SET>(bs, blocks, n, nullptr);
}

int lbm_blocks_get_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double* out) {
  return get_rows(bs, blocks, n, out, false);
}

int lbm_blocks_merge_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n,
                                         const double* neighbour) {
  return run_boundary_density<BD_MERGE>(bs, blocks, n, neighbour);
}

}  // extern "C"
